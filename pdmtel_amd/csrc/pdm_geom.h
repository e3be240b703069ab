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
