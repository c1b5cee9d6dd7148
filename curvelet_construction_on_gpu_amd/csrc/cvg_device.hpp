// Device-side building blocks shared by the fused curvelet kernel and the
// diagnostic kernels, so that what the diagnostics check is the arithmetic
// the pipeline runs.
#ifndef CVG_DEVICE_HPP
#define CVG_DEVICE_HPP

#include <hip/hip_runtime.h>

#include "cvg_math.hpp"

namespace cvg {

// Per-(anchor, neighbour) quantities of compute_curve_bundle
// (cpu_curvelet.hpp:646-656, 730-733) that do not depend on the grid cell.
struct PairRec {
    double DX, DY, DX2, DY2;
    double cT0m_DX, sT0m_DY, cT0p_DX, sT0p_DY;
};

// Per-edge trig slots: rows 0-5 are structure of arrays (stride = number of
// edges n), then one block of 2 Kt values per edge (edge-major, so the fused
// kernel reads an anchor's grid angles from one or two cache lines):
//   row 0 cos(theta)      row 1 sin(theta)
//   row 2 sin(theta+dt)   row 3 cos(theta+dt)   row 4 sin(theta-dt)   row 5 cos(theta-dt)
//   [6n + 2(i Kt + j)] = sin(theta_i + off_t[j]),  [6n + 2(i Kt + j) + 1] = cos(theta_i + off_t[j])
enum {
    TRIG_COS = 0,
    TRIG_SIN = 1,
    TRIG_SIN_P = 2,
    TRIG_COS_P = 3,
    TRIG_SIN_M = 4,
    TRIG_COS_M = 5,
    TRIG_GRID = 6
};

// Direction class of a neighbour: 1 when dot(To2Pi(atan2(ny-ay, nx-ax)), theta_a)
// > 0 (forward run), 2 when < 0 (backward run), 0 otherwise
// (cpu_curvelet.hpp:459-461, curvelet_utils.hpp:58-69).
//
// The reference's dot is cos(phi - theta_a), phi = atan2(vy, vx), v = n - a,
// up to its evaluation errors (atan2, To2Pi, cos, sin within an ulp each, the
// products and the sum), below 2e-15 in all.  The projection p = vx cos_at +
// vy sin_at, from the correctly rounded cos/sin of theta_a, is within 6e-16 |v|
// of |v| cos(phi - theta_a).  So when |p| > 1e-12 |v| the two signs agree and
// the class is sign(p); only near-perpendicular neighbours evaluate the
// reference's expression.
__device__ inline int direction_class(double ax, double ay, double cos_at, double sin_at, double nx, double ny,
                                      const double *sc_tab, const double *at_tab)
{
    const double vx = nx - ax, vy = ny - ay;
    const double p = vx * cos_at + vy * sin_at;
    if (fabs(p) > 1e-12 * sqrt(vx * vx + vy * vy)) return p > 0 ? 1 : 2;
    const double dir = wrap_two_pi(cr_atan2(vy, vx, at_tab));
    double sd, cd;
    cr_sincos(dir, sc_tab, &sd, &cd);
    const double dot = cd * cos_at + sd * sin_at;
    return dot > 0 ? 1 : (dot < 0 ? 2 : 0);
}

// The cell-independent part of one pair.
__device__ inline PairRec pair_record(double ax, double ay, double nx, double ny, double sT0p, double cT0p,
                                      double sT0m, double cT0m)
{
    PairRec r;
    r.DX = ax - nx;
    r.DY = ay - ny;
    r.DX2 = r.DX * r.DX;
    r.DY2 = r.DY * r.DY;
    r.cT0m_DX = r.DX * cT0m;
    r.sT0m_DY = r.DY * sT0m;
    r.cT0p_DX = r.DX * cT0p;
    r.sT0p_DY = r.DY * sT0p;
    return r;
}

// One grid cell of the circular-arc bundle transport (Prop. 4),
// compute_curve_bundle's loop body (cpu_curvelet.hpp:697-751 for branch A,
// :762-806 for branch B), starting from [-max_k, max_k].  Operand order and
// rounding points are the reference's; the updates use the reference's
// compare-and-assign, which ignores NaN candidates.
//   ox    = delta_x2 of the cell's row i
//   sT2t, cT2t = sin/cos(theta_a + delta_t2)
//   sm, sp = sin(T2t - (theta_n - dt)), sin(T2t - (theta_n + dt))
__device__ inline void bundle_cell(const PairRec &p, bool branch_a, double ox, double sT2t, double cT2t, double sm,
                                   double sp, double dx, double dx2, double max_k, double &lo, double &hi)
{
    const double tt = sT2t * p.DX - cT2t * p.DY;
    const double ttm = tt - ox;
    lo = -max_k;
    hi = max_k;
    if (branch_a) {
        const double den = -2.0 * ox * tt + p.DX2 + p.DY2 - dx2 + ox * ox;
        const double kx_min = 2.0 * (ttm - dx) / den;
        const double kx_max = 2.0 * (ttm + dx) / den;
        if (lo < kx_min) lo = kx_min;
        if (hi > kx_max) hi = kx_max;
        const double kt_min = sm / (-sm * ox + p.cT0m_DX + p.sT0m_DY);
        const double kt_max = sp / (-sp * ox + p.cT0p_DX + p.sT0p_DY);
        if (lo < kt_min) lo = kt_min;
        if (hi > kt_max) hi = kt_max;
    } else {
        const double den = -(2.0 * ox * tt - p.DX2 - p.DY2 + dx2 - ox * ox) / 2.0;
        const double kx_min = (ttm - dx) / den;
        const double kx_max = (ttm + dx) / den;
        if (lo < kx_min) lo = kx_min;
        if (hi > kx_max) hi = kx_max;
        const double kt_max = sm / (-sm * ox + p.cT0m_DX + p.sT0m_DY);
        if (hi > kt_max) hi = kt_max;
        const double kt_min = sp / (-sp * ox + p.cT0p_DX + p.sT0p_DY);
        if (lo < kt_min) lo = kt_min;
    }
}

}  // namespace cvg

#endif
