// Seeded synthetic third-order edge maps in the test_files format
// (x, y, theta, strength per edge).  The reference ships only three real edge
// files; these maps give the benchmark and the tests inputs of any size.
//
// Curves (circles/arcs, line segments, cubic Beziers) are sampled at 1 px arc
// length in curve order; positions get N(0, pos_sigma) noise, orientations are
// the tangent modulo pi plus N(0, theta_sigma) noise wrapped into [0, pi),
// strengths lognormal(2, 0.5).  T-junction patches add curves_per_patch line
// segments crossing one patch_size x patch_size box each.  Clutter edges
// (uniform position, uniform theta, strength lognormal(1, 0.5)) make up
// clutter_fraction of the result; samples outside the image are dropped.
// Randomness is xoshiro256** seeded through splitmix64, Gaussians by
// Box-Muller, so a (config, seed) pair always yields the same map on one
// platform (the transcendental calls are the host libm's).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/curvelet_gpu.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed)
    {
        uint64_t z = seed;
        for (auto &w : s) {
            z += 0x9e3779b97f4a7c15ULL;
            uint64_t t = z;
            t = (t ^ (t >> 30)) * 0xbf58476d1ce4e5b9ULL;
            t = (t ^ (t >> 27)) * 0x94d049bb133111ebULL;
            w = t ^ (t >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return static_cast<double>(next() >> 11) * 0x1p-53; }  // [0, 1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal()
    {
        double u1 = uniform();
        while (u1 <= 0.0) u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    }
};

struct Sample {
    double x, y, tangent;
};

double wrap_pi(double t)
{
    double a = std::fmod(t, kPi);
    if (a < 0) a += kPi;
    if (!(a >= 0 && a < kPi)) a = 0;
    return a;
}

void sample_polyline(const std::vector<Sample> &dense, std::vector<Sample> &out)
{
    // Resample a dense polyline at unit arc length.
    if (dense.size() < 2) return;
    double acc = 0.0, next_at = 0.0;
    for (size_t k = 1; k < dense.size(); ++k) {
        const double dx = dense[k].x - dense[k - 1].x, dy = dense[k].y - dense[k - 1].y;
        const double seg = std::sqrt(dx * dx + dy * dy);
        while (seg > 0 && next_at <= acc + seg) {
            const double f = (next_at - acc) / seg;
            out.push_back({dense[k - 1].x + f * dx, dense[k - 1].y + f * dy, std::atan2(dy, dx)});
            next_at += 1.0;
        }
        acc += seg;
    }
}

void add_circle(Rng &rng, const cvg_synth_config &c, std::vector<Sample> &out)
{
    const double cx = rng.uniform(0, c.width), cy = rng.uniform(0, c.height);
    double r;
    if (c.radius_log_uniform) {
        const double kmin = 1.0 / c.radius_max, kmax = 1.0 / c.radius_min;
        r = 1.0 / std::exp(rng.uniform(std::log(kmin), std::log(kmax)));
    } else {
        r = rng.uniform(c.radius_min, c.radius_max);
    }
    const double phi0 = rng.uniform(0, 2 * kPi);
    double extent = (rng.uniform() < 0.3) ? 2 * kPi : rng.uniform(kPi / 4, 2 * kPi);
    if (c.arc_max_length > 0 && r * extent > c.arc_max_length) extent = c.arc_max_length / r;
    const int n = static_cast<int>(std::floor(r * extent)) + 1;
    const double dir = (rng.uniform() < 0.5) ? 1.0 : -1.0;
    for (int k = 0; k < n; ++k) {
        const double phi = phi0 + dir * k / r;
        out.push_back({cx + r * std::cos(phi), cy + r * std::sin(phi), phi + dir * kPi / 2});
    }
}

void add_segment(Rng &rng, double cx, double cy, double len, std::vector<Sample> &out)
{
    const double a = rng.uniform(0, kPi);
    const double ux = std::cos(a), uy = std::sin(a);
    const int n = static_cast<int>(std::floor(len)) + 1;
    for (int k = 0; k < n; ++k) {
        const double s = k - len / 2;
        out.push_back({cx + s * ux, cy + s * uy, a});
    }
}

void add_bezier(Rng &rng, const cvg_synth_config &c, std::vector<Sample> &out)
{
    double px[4], py[4];
    if (c.bezier_extent > 0) {
        const double cx = rng.uniform(0, c.width), cy = rng.uniform(0, c.height);
        for (int k = 0; k < 4; ++k) {
            px[k] = cx + rng.uniform(-0.5, 0.5) * c.bezier_extent;
            py[k] = cy + rng.uniform(-0.5, 0.5) * c.bezier_extent;
        }
    } else {
        for (int k = 0; k < 4; ++k) {
            px[k] = rng.uniform(0, c.width);
            py[k] = rng.uniform(0, c.height);
        }
    }
    std::vector<Sample> dense;
    const int steps = 2048;
    for (int k = 0; k <= steps; ++k) {
        const double t = static_cast<double>(k) / steps, u = 1 - t;
        const double x = u * u * u * px[0] + 3 * u * u * t * px[1] + 3 * u * t * t * px[2] + t * t * t * px[3];
        const double y = u * u * u * py[0] + 3 * u * u * t * py[1] + 3 * u * t * t * py[2] + t * t * t * py[3];
        dense.push_back({x, y, 0.0});
    }
    sample_polyline(dense, out);
}

void emit_curve_samples(Rng &rng, const cvg_synth_config &c, const std::vector<Sample> &samples,
                        std::vector<double> &edges)
{
    const double ts = c.theta_sigma_deg * kPi / 180.0;
    for (const Sample &s : samples) {
        const double x = s.x + c.pos_sigma * rng.normal();
        const double y = s.y + c.pos_sigma * rng.normal();
        const double th = wrap_pi(s.tangent + ts * rng.normal());
        const double str = std::exp(2.0 + 0.5 * rng.normal());
        if (x < 0 || y < 0 || x >= c.width || y >= c.height) continue;
        edges.insert(edges.end(), {x, y, th, str});
    }
}

void set_defaults(cvg_synth_config *c)
{
    std::memset(c, 0, sizeof(*c));
    c->width = 320;
    c->height = 240;
    c->num_curves = 40;
    c->radius_min = 10;
    c->radius_max = 150;
    c->radius_log_uniform = 0;
    c->segment_min = 20;
    c->segment_max = 200;
    c->bezier_extent = 0;
    c->arc_max_length = 0;
    c->clutter_fraction = 0.1;
    c->junction_patches = 0;
    c->curves_per_patch = 20;
    c->patch_size = 50;
    c->pos_sigma = 0.1;
    c->theta_sigma_deg = 3.0;
}

}  // namespace

extern "C" {

int cvg_synth_preset(const char *name, cvg_synth_config *config)
{
    if (!name || !config) return CVG_ERR_INVALID_ARGUMENT;
    set_defaults(config);
    if (std::strcmp(name, "C1") == 0) {
        // 320x240, 40 curves, radius U[10,150], 10% clutter (~6k edges)
        config->arc_max_length = 150;
        return CVG_OK;
    }
    if (std::strcmp(name, "C2") == 0) {
        // 480x320, 150 curves, radius U[5,300], 20% clutter
        config->width = 480;
        config->height = 320;
        config->num_curves = 150;
        config->radius_min = 5;
        config->radius_max = 300;
        config->clutter_fraction = 0.2;
        config->bezier_extent = 160;
        config->arc_max_length = 200;
        return CVG_OK;
    }
    if (std::strcmp(name, "C3") == 0) {
        // 1920x1080, 1500 curves, curvature log-uniform [1/500, 1/3], 30% clutter,
        // T-junction patches of 20 crossing curves per 50x50 px
        config->width = 1920;
        config->height = 1080;
        config->num_curves = 1500;
        config->radius_min = 3;
        config->radius_max = 500;
        config->radius_log_uniform = 1;
        config->segment_max = 100;
        config->bezier_extent = 100;
        config->arc_max_length = 100;
        config->clutter_fraction = 0.3;
        config->junction_patches = 40;
        config->curves_per_patch = 20;
        config->patch_size = 50;
        return CVG_OK;
    }
    return CVG_ERR_INVALID_ARGUMENT;
}

int cvg_synth_generate(const cvg_synth_config *config, unsigned long long seed, double *edges, int capacity,
                       int *num_edges)
{
    if (!config || !num_edges || config->width <= 0 || config->height <= 0 || config->num_curves < 0 ||
        config->clutter_fraction < 0 || config->clutter_fraction >= 1)
        return CVG_ERR_INVALID_ARGUMENT;
    const cvg_synth_config &c = *config;
    Rng rng(seed);
    std::vector<double> out;
    std::vector<Sample> samples;
    for (int k = 0; k < c.num_curves; ++k) {
        samples.clear();
        const uint64_t kind = rng.next() % 3;
        if (kind == 0)
            add_circle(rng, c, samples);
        else if (kind == 1)
            add_segment(rng, rng.uniform(0, c.width), rng.uniform(0, c.height),
                        rng.uniform(c.segment_min, c.segment_max), samples);
        else
            add_bezier(rng, c, samples);
        emit_curve_samples(rng, c, samples, out);
    }
    for (int p = 0; p < c.junction_patches; ++p) {
        const double cx = rng.uniform(c.patch_size, c.width - c.patch_size);
        const double cy = rng.uniform(c.patch_size, c.height - c.patch_size);
        for (int k = 0; k < c.curves_per_patch; ++k) {
            samples.clear();
            const double ox = rng.uniform(-0.5, 0.5) * c.patch_size;
            const double oy = rng.uniform(-0.5, 0.5) * c.patch_size;
            add_segment(rng, cx + ox, cy + oy, c.patch_size, samples);
            emit_curve_samples(rng, c, samples, out);
        }
    }
    const long n_curve = static_cast<long>(out.size() / 4);
    const long n_clutter = std::lround(c.clutter_fraction / (1.0 - c.clutter_fraction) * n_curve);
    for (long k = 0; k < n_clutter; ++k) {
        const double x = rng.uniform(0, c.width), y = rng.uniform(0, c.height);
        const double th = rng.uniform(0, kPi);
        const double str = std::exp(1.0 + 0.5 * rng.normal());
        out.insert(out.end(), {x, y, th, str});
    }
    const long n = static_cast<long>(out.size() / 4);
    if (n > 0x7fffffffL) return CVG_ERR_UNSUPPORTED;
    *num_edges = static_cast<int>(n);
    if (edges) {
        const long m = n < capacity ? n : capacity;
        std::memcpy(edges, out.data(), static_cast<size_t>(m) * 4 * sizeof(double));
    }
    return CVG_OK;
}

}  // extern "C"
