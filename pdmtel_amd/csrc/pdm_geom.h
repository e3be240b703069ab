// pdm_geom.h -- bond predicate, beta factor and slit test, shared by host and device.
//
// Compiled with -ffp-contract=off everywhere: the bond set must be identical to
// the reference's, so the predicate keeps the reference's FP64 operation order
// (pdDataManager3.cpp:637-686 CalcDistanceToCube, 1573-1661 CalcDistanceToFace,
// 1663-1704 GetShortest/LongestBetween).  Device sqrt and division on gfx950
// are correctly rounded in FP64, as on the host.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

namespace pdm {

#define PDM_HD __host__ __device__ __forceinline__

PDM_HD double g_shortest(double x, double lo, double hi, double tol)
{
    if (x < lo - tol) return fabs(x - lo);
    else if (x > hi + tol) return fabs(x - hi);
    return 0.0;
}

PDM_HD double g_longest(double x, double lo, double hi, double tol)
{
    if (x < lo - tol) return fabs(x - hi);
    else if (x > hi + tol) return fabs(x - lo);
    if (x > 0.5 * (hi + lo)) return fabs(x - lo);
    return fabs(x - hi);
}

// Distance from point i to the face x_norm = x_face of node j's cell, found only when the
// line i->j pierces that face (within tol); DBL_MAX otherwise.  Only the minimum is needed.
PDM_HD double g_dist_to_face(int norm, double x_face, double x1i, double x2i, double x3i, double x1j,
                             double x2j, double x3j, double dx)
{
    double tol = 1e-4 * dx;
    int n1, n2, n3;
    if (norm == 1) { n1 = 1; n2 = n3 = 0; }
    else if (norm == 2) { n1 = n3 = 0; n2 = 1; }
    else { n1 = n2 = 0; n3 = 1; }
    if (fabs((x1i - x1j) * n1 + (x2i - x2j) * n2 + (x3i - x3j) * n3) < tol) return DBL_MAX;
    double x1lo = x1j - 0.5 * dx, x1hi = x1j + 0.5 * dx;
    double x2lo = x2j - 0.5 * dx, x2hi = x2j + 0.5 * dx;
    double x3lo = x3j - 0.5 * dx, x3hi = x3j + 0.5 * dx;
    double x_i = x1i * n1 + x2i * n2 + x3i * n3;
    double x_j = x1j * n1 + x2j * n2 + x3j * n3;
    double fac = (x_face - x_i) / (x_j - x_i);
    double x1int = x1i + fac * (x1j - x1i);
    double x2int = x2i + fac * (x2j - x2i);
    double x3int = x3i + fac * (x3j - x3i);
    if (x1int > x1lo - tol && x1int < x1hi + tol && x2int > x2lo - tol && x2int < x2hi + tol &&
        x3int > x3lo - tol && x3int < x3hi + tol) {
        double x1f = (1 - n1) * x1i + n1 * x_face;
        double x2f = (1 - n2) * x2i + n2 * x_face;
        double x3f = (1 - n3) * x3i + n3 * x_face;
        double x1s = (1 - n1) * g_shortest(x1f, x1lo, x1hi, tol);
        double x2s = (1 - n2) * g_shortest(x2f, x2lo, x2hi, tol);
        double x3s = (1 - n3) * g_shortest(x3f, x3lo, x3hi, tol);
        double ff_s = sqrt(x1s * x1s + x2s * x2s + x3s * x3s);
        double itf = sqrt((x1f - x1i) * (x1f - x1i) + (x2f - x2i) * (x2f - x2i) + (x3f - x3i) * (x3f - x3i));
        return sqrt(ff_s * ff_s + itf * itf);
    }
    return DBL_MAX;
}

// Shortest distance from node i to node j's cubic cell (CalcDistanceToCube's _min).
PDM_HD double g_dist_to_cube(double x1i, double x2i, double x3i, double x1j, double x2j, double x3j, double dx)
{
    double m = DBL_MAX, t;
    t = g_dist_to_face(1, x1j - 0.5 * dx, x1i, x2i, x3i, x1j, x2j, x3j, dx); m = (t < m) ? t : m;
    t = g_dist_to_face(1, x1j + 0.5 * dx, x1i, x2i, x3i, x1j, x2j, x3j, dx); m = (t < m) ? t : m;
    t = g_dist_to_face(2, x2j - 0.5 * dx, x1i, x2i, x3i, x1j, x2j, x3j, dx); m = (t < m) ? t : m;
    t = g_dist_to_face(2, x2j + 0.5 * dx, x1i, x2i, x3i, x1j, x2j, x3j, dx); m = (t < m) ? t : m;
    t = g_dist_to_face(3, x3j - 0.5 * dx, x1i, x2i, x3i, x1j, x2j, x3j, dx); m = (t < m) ? t : m;
    t = g_dist_to_face(3, x3j + 0.5 * dx, x1i, x2i, x3i, x1j, x2j, x3j, dx); m = (t < m) ? t : m;
    return m;
}

// Shortest and longest distances from node i to node j's cell, both as CalcDistanceToCube
// computes them (pdDataManager3.cpp:637-686): over the faces the line i->j pierces, the distance
// to the nearest and to the farthest point of the face (CalcDistanceToFace, :1573-1661).
PDM_HD void g_dist_to_cube_mm(double x1i, double x2i, double x3i, double x1j, double x2j, double x3j, double dx,
                              double* dmin, double* dmax)
{
    double mn = DBL_MAX, mx = DBL_MIN;
    const double tol = 1e-4 * dx;
    const double x1lo = x1j - 0.5 * dx, x1hi = x1j + 0.5 * dx;
    const double x2lo = x2j - 0.5 * dx, x2hi = x2j + 0.5 * dx;
    const double x3lo = x3j - 0.5 * dx, x3hi = x3j + 0.5 * dx;
    for (int f = 0; f < 6; ++f) {
        const int norm = f / 2 + 1;
        const int n1 = norm == 1, n2 = norm == 2, n3 = norm == 3;
        const double x_face = (f == 0) ? x1lo : (f == 1) ? x1hi : (f == 2) ? x2lo : (f == 3) ? x2hi : (f == 4) ? x3lo : x3hi;
        if (fabs((x1i - x1j) * n1 + (x2i - x2j) * n2 + (x3i - x3j) * n3) < tol) continue;
        double x_i = x1i * n1 + x2i * n2 + x3i * n3;
        double x_j = x1j * n1 + x2j * n2 + x3j * n3;
        double fac = (x_face - x_i) / (x_j - x_i);
        double x1int = x1i + fac * (x1j - x1i);
        double x2int = x2i + fac * (x2j - x2i);
        double x3int = x3i + fac * (x3j - x3i);
        if (!(x1int > x1lo - tol && x1int < x1hi + tol && x2int > x2lo - tol && x2int < x2hi + tol &&
              x3int > x3lo - tol && x3int < x3hi + tol))
            continue;
        double x1f = (1 - n1) * x1i + n1 * x_face;
        double x2f = (1 - n2) * x2i + n2 * x_face;
        double x3f = (1 - n3) * x3i + n3 * x_face;
        double x1s = (1 - n1) * g_shortest(x1f, x1lo, x1hi, tol);
        double x2s = (1 - n2) * g_shortest(x2f, x2lo, x2hi, tol);
        double x3s = (1 - n3) * g_shortest(x3f, x3lo, x3hi, tol);
        double x1l = (1 - n1) * g_longest(x1f, x1lo, x1hi, tol);
        double x2l = (1 - n2) * g_longest(x2f, x2lo, x2hi, tol);
        double x3l = (1 - n3) * g_longest(x3f, x3lo, x3hi, tol);
        double ff_s = sqrt(x1s * x1s + x2s * x2s + x3s * x3s);
        double ff_l = sqrt(x1l * x1l + x2l * x2l + x3l * x3l);
        double itf = sqrt((x1f - x1i) * (x1f - x1i) + (x2f - x2i) * (x2f - x2i) + (x3f - x3i) * (x3f - x3i));
        double a = sqrt(ff_s * ff_s + itf * itf), b = sqrt(ff_l * ff_l + itf * itf);
        mn = (a < mn) ? a : mn;
        mx = (b > mx) ? b : mx;
    }
    *dmin = mn;
    *dmax = mx;
}

// Adaptive-integration points of one family entry (CalcBondForceIntg, pdDataManager2.cpp:
// 1222-1291, with the nested trapezoidal rules of pdIntgManager.cpp:49-451): node j's cell,
// centre c = X_j - X_i in family coordinates, clipped to the horizon sphere around node i.
// Nothing when delta <= dmin; the whole cell when delta >= dmax (CalcConfigurationType3,
// :3471-3482); otherwise x (outer), y, z (inner) run between the cell faces and the sphere,
// whichever is nearer at the outer coordinates -- one rule for every configuration the
// reference enumerates in its 44 cases (:523-3469).  nseg = 2^(intg_point - 1) segments per
// direction, end points weighted 1/2; an interval of length <= 1e-10 adds nothing.  emit(x, y,
// z, w) receives each point; returns the count.
template <class EMIT>
PDM_HD int g_ai_cell_points(const double* c, double dmin, double dmax, double delta, double dx, int nseg, EMIT emit)
{
    if (delta <= dmin) return 0;
    const bool full = delta >= dmax;
    const double h = 0.5 * dx;
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = c[k] - h;
        hi[k] = c[k] + h;
    }
    double a1 = lo[0], b1 = hi[0];
    if (!full) {
        a1 = fmax(a1, -delta);
        b1 = fmin(b1, delta);
    }
    if (b1 - a1 <= 1.0e-10) return 0;
    const double H1 = (b1 - a1) / nseg;
    int cnt = 0;
    for (int ia = 0; ia <= nseg; ++ia) {
        const double x = ((nseg - ia) * a1 + ia * b1) / nseg;
        const double w1 = H1 * ((ia == 0 || ia == nseg) ? 0.5 : 1.0);
        double a2 = lo[1], b2 = hi[1];
        if (!full) {
            const double r1 = sqrt(fmax(delta * delta - x * x, 0.0));
            a2 = fmax(a2, -r1);
            b2 = fmin(b2, r1);
        }
        if (b2 - a2 <= 1.0e-10) continue;
        const double H2 = (b2 - a2) / nseg;
        for (int ib = 0; ib <= nseg; ++ib) {
            const double y = ((nseg - ib) * a2 + ib * b2) / nseg;
            const double w2 = H2 * ((ib == 0 || ib == nseg) ? 0.5 : 1.0);
            double a3 = lo[2], b3 = hi[2];
            if (!full) {
                const double r2 = sqrt(fmax(delta * delta - x * x - y * y, 0.0));
                a3 = fmax(a3, -r2);
                b3 = fmin(b3, r2);
            }
            if (b3 - a3 <= 1.0e-10) continue;
            const double H3 = (b3 - a3) / nseg;
            for (int ic = 0; ic <= nseg; ++ic) {
                const double z = ((nseg - ic) * a3 + ic * b3) / nseg;
                const double w3 = H3 * ((ic == 0 || ic == nseg) ? 0.5 : 1.0);
                emit(cnt, x, y, z, w1 * w2 * w3);
                ++cnt;
            }
        }
    }
    return cnt;
}

// InitBond's acceptance (pdDataManager2.cpp:33-45) for the pair (i, j), i the lower id.
// The coarse prefilter uses correctly rounded squares where the reference calls
// pow(x, 2.0); it cannot decide a pair, since a skipped pair has dmin > delta + 0.6 dx.
PDM_HD bool g_bonded(double x1i, double x2i, double x3i, double x1j, double x2j, double x3j, double dx,
                     double horizon, double* dist)
{
    double a = x1j - x1i, b = x2j - x2i, c = x3j - x3i;
    double d = sqrt(a * a + b * b + c * c);
    *dist = d;
    if (horizon < d - 1.5 * dx) return false;
    double dmin = g_dist_to_cube(x1i, x2i, x3i, x1j, x2j, x3j, dx);
    return dmin < horizon + 1.0e-4 * dx;
}

// CalcBetaFactor (pdDataManager3.cpp:1801-1831)
PDM_HD double g_beta(double dist, double delta, double dx, bool original_cci)
{
    double inner = delta - 0.5 * dx;
    double outer = delta + 0.5 * dx;
    if (original_cci) outer = delta;
    if (dist <= inner) return 1.0;
    else if (dist > inner && dist <= outer) return (delta + 0.5 * dx - dist) / dx;
    return 0.0;
}

// Slit cut (no reference implementation).  Type 1 keeps the axis-plane test that the
// reference has commented out (pdDataManager.cpp:1246-1290): the endpoints lie strictly
// on opposite sides and the crossing point lies in the rectangle, (q-lo)(q-hi) <= 0.
// Type 2 is an oriented rectangle (centre c, unit normal n, unit axes t and w = n x t).
PDM_HD bool g_slit_cut(int type, int axis, const double* p, const double* w, const double* xi, const double* xj)
{
    if (type == 1) {
        int d = axis;
        if (d < 0) return false;
        double xp = p[2 * d];
        if (!((xi[d] - xp) * (xj[d] - xp) < 0.)) return false;
        for (int o = 0; o < 3; ++o) {
            if (o == d) continue;
            double q = xi[o] + (xj[o] - xi[o]) * (xp - xi[d]) / (xj[d] - xi[d]);
            if (!((q - p[2 * o]) * (q - p[2 * o + 1]) <= 0.)) return false;
        }
        return true;
    }
    const double* c = p;
    const double* n = p + 3;
    const double* t = p + 6;
    double si = n[0] * (xi[0] - c[0]) + n[1] * (xi[1] - c[1]) + n[2] * (xi[2] - c[2]);
    double sj = n[0] * (xj[0] - c[0]) + n[1] * (xj[1] - c[1]) + n[2] * (xj[2] - c[2]);
    if (!(si * sj < 0.)) return false;
    double a = si / (si - sj);
    double h0 = xi[0] + a * (xj[0] - xi[0]) - c[0];
    double h1 = xi[1] + a * (xj[1] - xi[1]) - c[1];
    double h2 = xi[2] + a * (xj[2] - xi[2]) - c[2];
    double ht = h0 * t[0] + h1 * t[1] + h2 * t[2];
    double hw = h0 * w[0] + h1 * w[1] + h2 * w[2];
    return fabs(ht) <= p[9] && fabs(hw) <= p[10];
}

}  // namespace pdm
