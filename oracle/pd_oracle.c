/*
 * pd_oracle.c -- CPU oracle (checker) for the linear bond-based PD hot path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Never linked into or called by the product.
 *
 * A plain-C FP64 restatement of the reference's algorithms, written to keep
 * the reference's floating-point operation order (left-to-right products,
 * pow() where the reference calls pow(), no contraction: build with
 * -ffp-contract=off).  Node loops may run under OpenMP: every per-node
 * accumulation stays sequential over that node's family, so results are
 * identical for any thread count.
 *
 * Restated reference functions and what pins each one (tests/test_oracle.py):
 *   lattice      pdDataManager.cpp:846-909            closed-form id <-> (l,m,n) and coordinates
 *   regions      pdDataManager.cpp:913-1053           hand-built cases: last space wins, tolerance edges
 *   pair test    pdDataManager3.cpp:637-686,1573-1704 brute-force geometric distance point->cube on
 *                                                     random pairs; interior family 250 / 460;
 *                                                     reference bond totals (2,000-node bar: 168,392)
 *   bond search  pdDataManager2.cpp:12-70             O(N^2) reference loop == cell list, symmetry,
 *                                                     ascending families
 *   max family   pdDataManager.cpp:1345-1387          equals the interior family count (250 / 460)
 *   c, K, dt     pdDataManager.cpp:1314, pdMaterial.cpp:75-88, pdDataManager2.cpp:6-10
 *                                                     closed forms 12E/(pi d^4), sqrt(K/rho)
 *   beta         pdDataManager3.cpp:1801-1831         closed form, piecewise breakpoints
 *   bdryEff      pdDataManager.cpp:1340-1473, pdDataManager2.cpp:1353-1485
 *                                                     type-one closed form; full families -> 1;
 *                                                     energy branch vs an independent numpy sum
 *   CCI force    pdDataManager2.cpp:1293-1351         2-node closed form, eta=0 -> 0, rigid
 *                                                     translation -> 0, antisymmetry, sum F = 0
 *   RK4 + BC     pdDataManager.cpp:1767-2134,2497-2602, pdDataManager2.cpp:410-439
 *                                                     2-node closed-form recurrence written
 *                                                     independently in the test; uniform velocity
 *                                                     advances U by 0.75 V dt per step
 *   time factor  pdBdryConditionDispl.cpp:15-27, pdBdryConditionVeloc.cpp:15-18
 *                                                     closed form
 *   driver loop  pdSolve.cpp:55-120                   step count / shortened last step closed form
 *   slit cut     (no reference implementation; axis form follows the commented test at
 *                pdDataManager.cpp:1246-1290)        brute-force segment/rectangle test in numpy
 *   tangent K u  pdDataManager2.cpp:1763-1852 (operator form)
 *                                                     finite difference of the force (eta -> 0)
 *   Gauss rules  pdDataManager3.cpp:1706-1799         numpy.polynomial.legendre.leggauss; the
 *                                                     rule integrates x^(2n-1) exactly
 *   Gauss points pdDataManager2.cpp:72-104            closed-form positions x_j + (dx/2) a
 *   MLS          pdDataManager2.cpp:106-256,1547-1614 reproduces quadratic fields (the reference's
 *                                                     ConsistencyCheck criteria); equals an
 *                                                     independent numpy weighted least squares
 *   FGI force    pdDataManager.cpp:1918-1957          n = 1 equals the modified-CCI force;
 *                                                     eta = 0 -> 0; rigid translation -> 0; the
 *                                                     force converges in n to a numpy cubature
 *                                                     of the bond force over node j's cell
 *   Gauss-pt u   pdDataManager.cpp:2402-2495,2604-2703
 *                                                     affine fields interpolate exactly; BC
 *                                                     points take the prescribed value
 */
#include "pd_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MEGA DBL_MAX
#define TINY DBL_MIN

/* The reference squares with pow(x, 2.0).  glibc's pow is not correctly rounded
 * for y = 2 (about 1 in 1,200 random doubles differ from x*x by one ulp), so the
 * oracle can evaluate those calls either with glibc pow (mode 0, literal) or as
 * the correctly rounded square x*x (mode 1), which is what IEEE gives the device. */
static int g_square_exact = 0;
void orc_set_square_mode(int exact) { g_square_exact = exact; }
static double sq(double x) { return g_square_exact ? x * x : pow(x, 2.0); }

struct orc_state {
    orc_problem p;
    double pi;
    double tol;
    int64_t n;
    double* X;     /* 3N */
    int* mat;      /* N */
    int* bc;       /* N, -1 = none */
    int64_t* rp;   /* N+1 */
    int32_t* col;  /* E */
    uint8_t* cut;  /* E */
    double* be;    /* N bdryEff factor */
    double* spring;/* per material */
    double min_bond;
    double max_c;
    int max_family;
    double* trac;  /* 3N traction body-force density (extension) */
    /* fixed Gaussian integration (fgi_on): ng = n^3 points per node */
    int ng;
    double* gofs;  /* 3 ng: point offsets 0.5 dx a from the node */
    double* gw;    /* 3 ng: the three 1-D weights of each point */
    int64_t* cb_rp;/* N+1: MLS contributing nodes of each node (n > 1) */
    int32_t* cb;
    double* phi;   /* ng * cb_rp[N]: node i's block [g][k] at ng * cb_rp[i] */
    double* G;     /* 3 N ng: Gauss-point displacements */
    /* adaptive integration (ai_on): trapezoidal points of every entry, in family coordinates */
    int ai;
    int64_t* ai_ep;  /* E+1: points of each entry (reference CSR order) */
    double* ai_xw;   /* 4 P: x, y, z relative to node i, weight */
    int64_t* ai_po;  /* P+1: MLS coefficients of each point over node j's contributing nodes */
    double* ai_phi;
    /* dynamic state */
    double *U, *V, *Vb, *elast;
    double t, dt_cur, dt_stable, ke, ee;
    int steps_done, last_step;
};

/* ------------------------------------------------------------------ */
/* geometry helpers: pdDataManager3.cpp:1663-1704                       */
static double shortest_between(double x, double lo, double hi, double tol)
{
    if (x < lo - tol) return fabs(x - lo);
    else if (x > hi + tol) return fabs(x - hi);
    return 0.0;
}

static double longest_between(double x, double lo, double hi, double tol)
{
    if (x < lo - tol) return fabs(x - hi);
    else if (x > hi + tol) return fabs(x - lo);
    if (x > 0.5 * (hi + lo)) return fabs(x - lo);
    return fabs(x - hi);
}

/* pdDataManager3.cpp:1573-1661, with positions X+U (U = 0 at bond time). */
static void dist_to_face(int norm, double x_face, const double* pi_, const double* pj_, double dx,
                         double* dmin, double* dmax)
{
    double tol = 1e-4 * dx;
    int n1, n2, n3;
    if (norm == 1) { n1 = 1; n2 = n3 = 0; }
    else if (norm == 2) { n1 = n3 = 0; n2 = 1; }
    else { n1 = n2 = 0; n3 = 1; }
    double x1i = pi_[0], x2i = pi_[1], x3i = pi_[2];
    double x1j = pj_[0], x2j = pj_[1], x3j = pj_[2];
    if (fabs((x1i - x1j) * n1 + (x2i - x2j) * n2 + (x3i - x3j) * n3) < tol) {
        *dmin = MEGA; *dmax = TINY; return;
    }
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
        double x1s = (1 - n1) * shortest_between(x1f, x1lo, x1hi, tol);
        double x2s = (1 - n2) * shortest_between(x2f, x2lo, x2hi, tol);
        double x3s = (1 - n3) * shortest_between(x3f, x3lo, x3hi, tol);
        double x1l = (1 - n1) * longest_between(x1f, x1lo, x1hi, tol);
        double x2l = (1 - n2) * longest_between(x2f, x2lo, x2hi, tol);
        double x3l = (1 - n3) * longest_between(x3f, x3lo, x3hi, tol);
        double ff_s = sqrt(x1s * x1s + x2s * x2s + x3s * x3s);
        double ff_l = sqrt(x1l * x1l + x2l * x2l + x3l * x3l);
        double itf = sqrt((x1f - x1i) * (x1f - x1i) + (x2f - x2i) * (x2f - x2i) + (x3f - x3i) * (x3f - x3i));
        *dmin = sqrt(ff_s * ff_s + itf * itf);
        *dmax = sqrt(ff_l * ff_l + itf * itf);
        return;
    }
    *dmin = MEGA; *dmax = TINY;
}

/* pdDataManager3.cpp:637-686 */
static void dist_to_cube(const double* xi, const double* xj, double dx, double* mn, double* mx)
{
    double a = MEGA, b = TINY, t0, t1;
    double lo[3] = {xj[0] - 0.5 * dx, xj[1] - 0.5 * dx, xj[2] - 0.5 * dx};
    double hi[3] = {xj[0] + 0.5 * dx, xj[1] + 0.5 * dx, xj[2] + 0.5 * dx};
    for (int d = 0; d < 3; ++d) {
        dist_to_face(d + 1, lo[d], xi, xj, dx, &t0, &t1);
        a = (t0 < a) ? t0 : a; b = (b < t1) ? t1 : b;
        dist_to_face(d + 1, hi[d], xi, xj, dx, &t0, &t1);
        a = (t0 < a) ? t0 : a; b = (b < t1) ? t1 : b;
    }
    *mn = a; *mx = b;
}

int orc_pair_bonded(const orc_problem* p, const double* xi, const double* xj, double* dmin, double* dmax)
{
    /* pdDataManager2.cpp:33-45: prefilter then dmin < delta + 1e-4 dx */
    double dx = p->dx;
    double d = sqrt(sq(xj[0] - xi[0]) + sq(xj[1] - xi[1]) + sq(xj[2] - xi[2]));
    if (p->horizon < d - 1.5 * dx) { *dmin = MEGA; *dmax = TINY; return 0; }
    dist_to_cube(xi, xj, dx, dmin, dmax);
    return *dmin < p->horizon + 1.0e-4 * dx;
}

/* Slit cut test (extension).  Type 1 keeps the reference's commented axis-plane
 * test (pdDataManager.cpp:1246-1290); type 2 is an oriented rectangle. */
static void slit_frame(const orc_slit* s, double* n, double* t, double* w)
{
    double nn = sqrt(s->p[3] * s->p[3] + s->p[4] * s->p[4] + s->p[5] * s->p[5]);
    n[0] = s->p[3] / nn; n[1] = s->p[4] / nn; n[2] = s->p[5] / nn;
    double tn = s->p[6] * n[0] + s->p[7] * n[1] + s->p[8] * n[2];
    double tt[3] = {s->p[6] - tn * n[0], s->p[7] - tn * n[1], s->p[8] - tn * n[2]};
    double tl = sqrt(tt[0] * tt[0] + tt[1] * tt[1] + tt[2] * tt[2]);
    t[0] = tt[0] / tl; t[1] = tt[1] / tl; t[2] = tt[2] / tl;
    w[0] = n[1] * t[2] - n[2] * t[1];
    w[1] = n[2] * t[0] - n[0] * t[2];
    w[2] = n[0] * t[1] - n[1] * t[0];
}

static int slit_cuts(const orc_slit* s, const double* xi, const double* xj)
{
    if (s->type == 1) {
        const double* b = s->p;
        int d = -1;
        for (int k = 0; k < 3; ++k) if (b[2 * k] == b[2 * k + 1]) { d = k; break; }
        if (d < 0) return 0;
        double xp = b[2 * d];
        if (!((xi[d] - xp) * (xj[d] - xp) < 0.)) return 0;
        for (int o = 0; o < 3; ++o) {
            if (o == d) continue;
            double q = xi[o] + (xj[o] - xi[o]) * (xp - xi[d]) / (xj[d] - xi[d]);
            if (!((q - b[2 * o]) * (q - b[2 * o + 1]) <= 0.)) return 0;
        }
        return 1;
    }
    double n[3], t[3], w[3];
    slit_frame(s, n, t, w);
    const double* c = s->p;
    double si = n[0] * (xi[0] - c[0]) + n[1] * (xi[1] - c[1]) + n[2] * (xi[2] - c[2]);
    double sj = n[0] * (xj[0] - c[0]) + n[1] * (xj[1] - c[1]) + n[2] * (xj[2] - c[2]);
    if (!(si * sj < 0.)) return 0;
    double a = si / (si - sj);
    double h[3] = {xi[0] + a * (xj[0] - xi[0]) - c[0], xi[1] + a * (xj[1] - xi[1]) - c[1],
                   xi[2] + a * (xj[2] - xi[2]) - c[2]};
    double ht = h[0] * t[0] + h[1] * t[1] + h[2] * t[2];
    double hw = h[0] * w[0] + h[1] * w[1] + h[2] * w[2];
    return fabs(ht) <= s->p[9] && fabs(hw) <= s->p[10];
}

int orc_pair_cut(const orc_problem* p, const double* xi, const double* xj)
{
    for (int k = 0; k < p->nslit; ++k)
        if (slit_cuts(&p->slit[k], xi, xj)) return 1;
    return 0;
}

/* pdDataManager3.cpp:1801-1831 */
double orc_beta(double dist, double delta, double dx, int cci_on, int cci_mdf)
{
    double inner = delta - 0.5 * dx;
    double outer = delta + 0.5 * dx;
    if (cci_on && !cci_mdf) outer = delta;
    if (dist <= inner) return 1.0;
    else if (dist > inner && dist <= outer) return (delta + 0.5 * dx - dist) / dx;
    return 0.0;
}

/* pdDataManager2.cpp:1293-1351 (array overload used by the RK4 path) */
void orc_bond_force_cci(const orc_problem* p, const double* xi, const double* eta, double ecrit,
                        double spring, double bij, double effWeight, double* f, double* elast)
{
    double xi1 = xi[0], xi2 = xi[1], xi3 = xi[2];
    double e1 = eta[0], e2 = eta[1], e3 = eta[2];
    double r = sqrt(xi1 * xi1 + xi2 * xi2 + xi3 * xi3);
    double pp = sqrt(sq(xi1 + e1) + sq(xi2 + e2) + sq(xi3 + e3));
    double dxe = effWeight * p->dx / 2.0;
    double beta = orc_beta(r, p->horizon, dxe, p->cci_on, p->cci_mdf);
    double u = pp - r;
    double st = u / r;
    if (u <= 0.0 || (u > 0.0 && st <= ecrit)) {
        if (p->cci_on) {
            double vol = p->dx * p->dx * p->dx;
            f[0] = bij * vol * beta * spring * st * (xi1 + e1) / pp;
            f[1] = bij * vol * beta * spring * st * (xi2 + e2) / pp;
            f[2] = bij * vol * beta * spring * st * (xi3 + e3) / pp;
        } else {
            f[0] = bij * beta * spring * st * (xi1 + e1) / pp;
            f[1] = bij * beta * spring * st * (xi2 + e2) / pp;
            f[2] = bij * beta * spring * st * (xi3 + e3) / pp;
        }
        *elast = bij * 0.5 * (0.5 * beta * (spring / r) * u * u) * pow(0.5 * effWeight, 3);
    } else {
        f[0] = f[1] = f[2] = 0.0;
        *elast = 0.0;
    }
}

/* pdBdryConditionDispl.cpp:15-27 (types 1,3,4 share it); pdBdryConditionVeloc.cpp:15-18 */
double orc_time_factor(const orc_bc* bc, double t)
{
    if (bc->type == 2) return (t <= bc->t_end) ? 1.0 : 0.0;
    if (bc->load == 1.0) return (t <= bc->t_end) ? t / bc->t_end : 1.0;
    return 1.0;
}

/* ------------------------------------------------------------------ */
static uint64_t splitmix64(uint64_t* s)
{
    uint64_t z = (*s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static int in_space(const orc_space* sp, const double* x, double tol)
{
    const double* b = sp->b;
    if (sp->geo == 1)
        return x[0] >= b[0] - tol && x[0] <= b[1] + tol && x[1] >= b[2] - tol && x[1] <= b[3] + tol &&
               x[2] >= b[4] - tol && x[2] <= b[5] + tol;
    if (sp->geo == 2) {
        double dist = sqrt((x[0] - b[1]) * (x[0] - b[1]) + (x[1] - b[2]) * (x[1] - b[2]));
        return dist <= b[0] + tol && x[2] >= b[3] - tol && x[2] <= b[4] + tol;
    }
    return 0;
}

orc_state* orc_create(const orc_problem* p)
{
    orc_state* s = (orc_state*)calloc(1, sizeof(orc_state));
    s->p = *p;
    s->pi = acos(-1.0);
    s->tol = 1.0e-4 * p->dx;
    return s;
}

void orc_destroy(orc_state* s)
{
    if (!s) return;
    free(s->X); free(s->mat); free(s->bc); free(s->rp); free(s->col); free(s->cut);
    free(s->be); free(s->spring); free(s->trac);
    free(s->U); free(s->V); free(s->Vb); free(s->elast);
    free(s->gofs); free(s->gw); free(s->cb_rp); free(s->cb); free(s->phi); free(s->G);
    free(s->ai_ep); free(s->ai_xw); free(s->ai_po); free(s->ai_phi);
    free(s);
}

/* InitNode, pdDataManager.cpp:813-1098 (lattice branch), plus the seeded jitter extension */
static int init_nodes(orc_state* s)
{
    const orc_problem* p = &s->p;
    int64_t cap = (int64_t)p->nx * p->ny * p->nz;
    s->X = (double*)malloc(sizeof(double) * 3 * (cap > 0 ? cap : 1));
    int64_t id = 0;
    double tol = s->tol;
    for (int n = 1; n != p->nz + 1; ++n)
        for (int m = 1; m != p->ny + 1; ++m)
            for (int l = 1; l != p->nx + 1; ++l) {
                double x[3];
                x[0] = (l - (p->nx + 1) * 0.5) * p->dx + p->cx;
                x[1] = (m - (p->ny + 1) * 0.5) * p->dx + p->cy;
                x[2] = (n - (p->nz + 1) * 0.5) * p->dx + p->cz;
                int del = 0;
                for (int d = 0; d < p->nds; ++d)
                    if (in_space(&p->ds[d], x, tol)) del = 1;
                if (!del) {
                    memcpy(&s->X[3 * id], x, sizeof x);
                    ++id;
                }
            }
    s->n = id;
    if (p->jitter != 0.0) {
        uint64_t st = p->seed;
        for (int64_t i = 0; i < s->n; ++i)
            for (int c = 0; c < 3; ++c) {
                double u01 = (double)(splitmix64(&st) >> 11) * 0x1.0p-53;
                s->X[3 * i + c] += (2.0 * u01 - 1.0) * p->jitter * p->dx;
            }
    }
    s->mat = (int*)malloc(sizeof(int) * (s->n ? s->n : 1));
    s->bc = (int*)malloc(sizeof(int) * (s->n ? s->n : 1));
    for (int64_t i = 0; i < s->n; ++i) {
        const double* x = &s->X[3 * i];
        int m = -1;
        /* every material space overwrites; a node outside a later space loses its
         * material (pdDataManager.cpp:918-968), so effectively the last space decides */
        for (int k = 0; k < p->nms; ++k) m = in_space(&p->ms[k], x, tol) ? p->ms[k].ref : -1;
        if (p->nms == 0 || m < 0) return -3; /* node not in any material space (fatal) */
        s->mat[i] = m;
        int b = -1;
        for (int k = 0; k < p->nbs; ++k)
            if (in_space(&p->bs[k], x, tol)) b = p->bs[k].ref;
        s->bc[i] = b;
    }
    return 0;
}

typedef struct { int64_t cnt, cap; int32_t* v; } ivec;
static void ivec_push(ivec* a, int32_t x)
{
    if (a->cnt == a->cap) { a->cap = a->cap ? 2 * a->cap : 64; a->v = (int32_t*)realloc(a->v, sizeof(int32_t) * a->cap); }
    a->v[a->cnt++] = x;
}
static int cmp_i32(const void* a, const void* b)
{
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

/* InitBond, pdDataManager2.cpp:12-70 */
/* Cell list with edge >= delta + 1.5 dx: every pair the reference can accept lies in
 * the 27 cells around a node's own. */
typedef struct {
    double lo[3], h;
    int64_t nc[3];
    int64_t* head;   /* ncell + 1 */
    int64_t* cid;    /* n */
    int64_t* order;  /* n, nodes sorted by cell */
} cell_grid;

static void grid_build(const orc_state* s, cell_grid* g)
{
    const orc_problem* p = &s->p;
    int64_t n = s->n;
    double hi[3] = {-MEGA, -MEGA, -MEGA};
    for (int c = 0; c < 3; ++c) g->lo[c] = MEGA;
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            double v = s->X[3 * i + c];
            g->lo[c] = v < g->lo[c] ? v : g->lo[c];
            hi[c] = v > hi[c] ? v : hi[c];
        }
    g->h = p->horizon + 1.5 * p->dx;
    for (int c = 0; c < 3; ++c) g->nc[c] = (n ? (int64_t)floor((hi[c] - g->lo[c]) / g->h) : 0) + 1;
    int64_t ncell = g->nc[0] * g->nc[1] * g->nc[2];
    g->head = (int64_t*)calloc(ncell + 1, sizeof(int64_t));
    g->cid = (int64_t*)malloc(sizeof(int64_t) * (n ? n : 1));
    g->order = (int64_t*)malloc(sizeof(int64_t) * (n ? n : 1));
    for (int64_t i = 0; i < n; ++i) {
        int64_t q[3];
        for (int c = 0; c < 3; ++c) {
            q[c] = (int64_t)floor((s->X[3 * i + c] - g->lo[c]) / g->h);
            if (q[c] >= g->nc[c]) q[c] = g->nc[c] - 1;
        }
        g->cid[i] = q[0] + g->nc[0] * (q[1] + g->nc[1] * q[2]);
        g->head[g->cid[i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; ++c) g->head[c + 1] += g->head[c];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (ncell + 1));
    memcpy(fill, g->head, sizeof(int64_t) * (ncell + 1));
    for (int64_t i = 0; i < n; ++i) g->order[fill[g->cid[i]]++] = i;
    free(fill);
}

static void grid_free(cell_grid* g)
{
    free(g->head);
    free(g->cid);
    free(g->order);
}

/* The family of node i (ascending ids): every j != i for which the reference's pair test,
 * evaluated on (lower id, higher id) as InitBond's i < j loop does, accepts the bond.
 * *minb takes the minimum bond length over the accepted pairs with j > i. */
static void gather_family(const orc_state* s, const cell_grid* g, int64_t i, ivec* out, double* minb)
{
    const orc_problem* p = &s->p;
    int64_t c0 = g->cid[i];
    int64_t q0 = c0 % g->nc[0], q1 = (c0 / g->nc[0]) % g->nc[1], q2 = c0 / (g->nc[0] * g->nc[1]);
    for (int64_t a = q2 - 1; a <= q2 + 1; ++a) {
        if (a < 0 || a >= g->nc[2]) continue;
        for (int64_t b = q1 - 1; b <= q1 + 1; ++b) {
            if (b < 0 || b >= g->nc[1]) continue;
            for (int64_t c = q0 - 1; c <= q0 + 1; ++c) {
                if (c < 0 || c >= g->nc[0]) continue;
                int64_t cc = c + g->nc[0] * (b + g->nc[1] * a);
                for (int64_t k = g->head[cc]; k < g->head[cc + 1]; ++k) {
                    int64_t j = g->order[k];
                    if (j == i) continue;
                    int64_t lo_id = i < j ? i : j, hi_id = i < j ? j : i;
                    double dmin, dmax;
                    const double* xa = &s->X[3 * lo_id];
                    const double* xb = &s->X[3 * hi_id];
                    if (orc_pair_bonded(p, xa, xb, &dmin, &dmax)) {
                        if (j > i && minb) {
                            double d = sqrt(sq(xb[0] - xa[0]) + sq(xb[1] - xa[1]) + sq(xb[2] - xa[2]));
                            *minb = (d < *minb) ? d : *minb;
                        }
                        ivec_push(out, (int32_t)j);
                    }
                }
            }
        }
    }
    qsort(out->v, out->cnt, sizeof(int32_t), cmp_i32);
}

/* InitBond, pdDataManager2.cpp:12-70 */
static int init_bonds(orc_state* s, int method, int nthreads)
{
    const orc_problem* p = &s->p;
    int64_t n = s->n;
    ivec* fam = (ivec*)calloc(n ? n : 1, sizeof(ivec));
    double minb = p->horizon;
    if (method == 0) {
        for (int64_t i = 0; i < n; ++i)
            for (int64_t j = i + 1; j < n; ++j) {
                double dmin, dmax;
                if (orc_pair_bonded(p, &s->X[3 * i], &s->X[3 * j], &dmin, &dmax)) {
                    const double* a = &s->X[3 * i]; const double* b = &s->X[3 * j];
                    double d = sqrt(sq(b[0] - a[0]) + sq(b[1] - a[1]) + sq(b[2] - a[2]));
                    minb = (d < minb) ? d : minb;
                    ivec_push(&fam[i], (int32_t)j);
                    ivec_push(&fam[j], (int32_t)i);
                }
            }
    } else {
        /* each node's family is gathered by that node alone, so the node loop runs in
         * parallel; every pair is still tested as (lower id, higher id) */
        cell_grid g;
        grid_build(s, &g);
#ifdef _OPENMP
#pragma omp parallel if (nthreads != 1)
#endif
        {
            double mloc = minb;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 256)
#endif
            for (int64_t i = 0; i < n; ++i) gather_family(s, &g, i, &fam[i], &mloc);
#ifdef _OPENMP
#pragma omp critical
#endif
            minb = (mloc < minb) ? mloc : minb;
        }
        grid_free(&g);
    }
    s->min_bond = minb;
    s->rp = (int64_t*)malloc(sizeof(int64_t) * (n + 1));
    s->rp[0] = 0;
    for (int64_t i = 0; i < n; ++i) s->rp[i + 1] = s->rp[i] + fam[i].cnt;
    int64_t E = s->rp[n];
    s->col = (int32_t*)malloc(sizeof(int32_t) * (E ? E : 1));
    s->cut = (uint8_t*)malloc(E ? E : 1);
    for (int64_t i = 0; i < n; ++i) {
        if (fam[i].cnt) memcpy(&s->col[s->rp[i]], fam[i].v, sizeof(int32_t) * fam[i].cnt);
        free(fam[i].v);
    }
    free(fam);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256) if (nthreads != 1)
#endif
    for (int64_t i = 0; i < n; ++i)
        for (int64_t e = s->rp[i]; e < s->rp[i + 1]; ++e) {
            int64_t j = s->col[e];
            int64_t a = i < j ? i : j, b = i < j ? j : i;
            s->cut[e] = (uint8_t)orc_pair_cut(p, &s->X[3 * a], &s->X[3 * b]);
        }
    return 0;
}

/* InitMaterial, pdDataManager.cpp:1297-1338 */
static void init_material(orc_state* s)
{
    const orc_problem* p = &s->p;
    s->spring = (double*)malloc(sizeof(double) * (p->nmat ? p->nmat : 1));
    s->max_c = 0.0;
    for (int m = 0; m < p->nmat; ++m) {
        double nu = (p->mat[m].type == 1) ? 0.25 : p->mat[m].nu;
        double K = p->mat[m].E / (3.0 * (1.0 - 2.0 * nu));
        s->spring[m] = (p->mat[m].type == 1) ? 18 * K / (s->pi * pow(p->horizon, 4.0)) : 0.0;
        double c = sqrt(K / p->mat[m].rho);
        s->max_c = (s->max_c < c) ? c : s->max_c;
    }
}

double orc_spring(const orc_state* s, int m) { return s->spring[m]; }

/* CalcBdryEffFactor, pdDataManager2.cpp:1353-1477 */
static double bdryeff_energy(const orc_state* s, int64_t i, const int32_t* fam, int64_t cnt)
{
    const orc_problem* p = &s->p;
    double e11 = 1.0e-4;
    const orc_mat* mi = &p->mat[s->mat[i]];
    double ecrit = mi->ecrit, spring = s->spring[s->mat[i]];
    double eps11 = e11, eps22 = e11, eps33 = e11;
    double eps12 = 0.0, eps13 = 0.0, eps21 = 0.0, eps23 = 0.0, eps31 = 0.0, eps32 = 0.0;
    double vol = pow(p->dx, 3.0);
    double sum = 0.0;
    const double* xi_ = &s->X[3 * i];
    /* slit cuts do not enter: the factor is that of the intact family (a slit is a
     * pre-crack introduced after the surface calibration) */
    for (int64_t e = 0; e < cnt; ++e) {
        const double* xj = &s->X[3 * (int64_t)fam[e]];
        double xi1 = xj[0] - xi_[0], xi2 = xj[1] - xi_[1], xi3 = xj[2] - xi_[2];
        double eta1 = eps11 * xi1 + eps12 * xi2 + eps13 * xi3;
        double eta2 = eps21 * xi1 + eps22 * xi2 + eps23 * xi3;
        double eta3 = eps31 * xi1 + eps32 * xi2 + eps33 * xi3;
        double r = sqrt(xi1 * xi1 + xi2 * xi2 + xi3 * xi3);
        double p1 = xi1 + eta1, p2 = xi2 + eta2, p3 = xi3 + eta3;
        double pp = sqrt(p1 * p1 + p2 * p2 + p3 * p3);
        double u = pp - r;
        double st = u / r;
        double el;
        if (u <= 0 || (u > 0 && st <= ecrit)) el = 0.5 * (spring / r) * u * u;
        else el = 0.0;
        double beta = orc_beta(r, p->horizon, p->dx, p->cci_on, p->cci_mdf);
        double elj = el * beta * vol;
        sum += 0.5 * elj;
    }
    double p_inf = 3.0 * mi->E * e11 * e11;
    return p_inf / sum;
}

/* CalcBdryEffFactorTypeOne, pdDataManager2.cpp:1479-1485 */
static double bdryeff_type_one(const orc_state* s, double h, int m)
{
    const orc_problem* p = &s->p;
    double full = 3.0 * p->mat[m].E;
    double part = full - 0.5 * s->spring[m] * s->pi *
                             (pow(p->horizon, 4) / 4.0 - h * pow(p->horizon, 3) / 3.0 + pow(h, 4) / 12.0);
    return full / part;
}

/* The largest family, counted on a virtual lattice around the origin
 * (pdDataManager.cpp:1345-1387). */
static int virtual_max_family(const orc_state* s)
{
    const orc_problem* p = &s->p;
    int vd = 2 * ((int)(p->horizon / p->dx) + 1);
    double o[3] = {0.0, 0.0, 0.0};
    int maxf = 0;
    for (int nn = 1; nn != vd + 1; ++nn)
        for (int m = 1; m != vd + 1; ++m)
            for (int l = 1; l != vd + 1; ++l) {
                double xj[3] = {(l - vd * 0.5) * p->dx, (m - vd * 0.5) * p->dx, (nn - vd * 0.5) * p->dx};
                double dmin, dmax;
                dist_to_cube(o, xj, p->dx, &dmin, &dmax);
                if (dmin < p->horizon + s->tol) ++maxf;
            }
    return maxf;
}

/* The boundary-effect factor of node i with family fam[0..cnt) (pdDataManager.cpp:1389-1473):
 * 1 for a full family, the type-one closed form in a single-surface band, else the energy
 * ratio. */
static double bdryeff_node(const orc_state* s, int64_t i, const int32_t* fam, int64_t cnt, int maxf)
{
    const orc_problem* p = &s->p;
    double x1L = -p->nx * 0.5 * p->dx + p->cx, x1H = p->nx * 0.5 * p->dx + p->cx;
    double x2L = -p->ny * 0.5 * p->dx + p->cy, x2H = p->ny * 0.5 * p->dx + p->cy;
    double x3L = -p->nz * 0.5 * p->dx + p->cz, x3H = p->nz * 0.5 * p->dx + p->cz;
    double ns = ((int)(p->horizon / p->dx)) * p->dx;
    double x1 = s->X[3 * i], x2 = s->X[3 * i + 1], x3 = s->X[3 * i + 2];
    int m = s->mat[i];
    if (cnt == maxf) return 1.0;
    if (x1 >= x1H - ns && x1 <= x1H && x2 >= x2L + ns && x2 <= x2H - ns && x3 >= x3L + ns && x3 <= x3H - ns)
        return bdryeff_type_one(s, x1H - x1, m);
    if (x1 >= x1L && x1 <= x1L + ns && x2 >= x2L + ns && x2 <= x2H - ns && x3 >= x3L + ns && x3 <= x3H - ns)
        return bdryeff_type_one(s, x1 - x1L, m);
    if (x2 >= x2H - ns && x2 <= x2H && x1 >= x1L + ns && x1 <= x1H - ns && x3 >= x3L + ns && x3 <= x3H - ns)
        return bdryeff_type_one(s, x2H - x2, m);
    if (x2 >= x2L && x2 <= x2L + ns && x1 >= x1L + ns && x1 <= x1H - ns && x3 >= x3L + ns && x3 <= x3H - ns)
        return bdryeff_type_one(s, x2 - x2L, m);
    if (x3 >= x3H - ns && x3 <= x3H && x1 >= x1L + ns && x1 <= x1H - ns && x2 >= x2L + ns && x2 <= x2H - ns)
        return bdryeff_type_one(s, x3H - x3, m);
    if (x3 >= x3L && x3 <= x3L + ns && x1 >= x1L + ns && x1 <= x1H - ns && x2 >= x2L + ns && x2 <= x2H - ns)
        return bdryeff_type_one(s, x3 - x3L, m);
    return bdryeff_energy(s, i, fam, cnt);
}

/* InitBdryEffectFactor, pdDataManager.cpp:1340-1473 */
static void init_bdryeff(orc_state* s, int nthreads)
{
    const orc_problem* p = &s->p;
    int64_t n = s->n;
    s->be = (double*)malloc(sizeof(double) * (n ? n : 1));
    for (int64_t i = 0; i < n; ++i) s->be[i] = 1.0;
    s->max_family = virtual_max_family(s);
    if (!p->bdryeff_on) return;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256) if (nthreads != 1)
#endif
    for (int64_t i = 0; i < n; ++i)
        s->be[i] = bdryeff_node(s, i, &s->col[s->rp[i]], s->rp[i + 1] - s->rp[i], s->max_family);
}

/* Traction as body force (extension): a TRACTION bc's value is a total force,
 * spread evenly over the nodes carrying that bc, as a force density per volume. */
static void init_traction(orc_state* s)
{
    const orc_problem* p = &s->p;
    int64_t n = s->n;
    s->trac = (double*)calloc(3 * (n ? n : 1), sizeof(double));
    if (!p->traction_body) return;
    int64_t cnt[ORC_MAX_BC] = {0};
    for (int64_t i = 0; i < n; ++i) if (s->bc[i] >= 0) cnt[s->bc[i]]++;
    double vol = pow(p->dx, 3.0);
    for (int64_t i = 0; i < n; ++i) {
        int b = s->bc[i];
        if (b < 0 || p->bc[b].type != 4) continue;
        for (int c = 0; c < 3; ++c)
            if (p->bc[b].f[c] == 1) s->trac[3 * i + c] = p->bc[b].v[c] / (double)cnt[b] / vol;
    }
}

/* ------------------------------------------------------------------ */
/* fixed Gaussian integration                                           */

/* Gauss-Legendre rule of order n on [-1, 1] (the rule SetGaussianPointsValue tabulates,
 * pdDataManager3.cpp:1706-1799), found by Newton's method on the three-term recurrence
 * (n+1) P_{n+1} = (2n+1) x P_n - n P_{n-1}, with w = 2 / ((1 - x^2) P_n'(x)^2).  Points are
 * listed in the reference's order: the zero (odd n), then each positive root in increasing
 * order followed by its negative; for n = 8 the four positive roots in increasing order,
 * then their negatives.  n = 7 and n > 8 have no rule in the reference. */
static void legendre(int n, double x, double* pn, double* dpn)
{
    double p0 = 1.0, p1 = x;
    for (int q = 1; q < n; ++q) {
        double p2 = ((2.0 * q + 1.0) * x * p1 - q * p0) / (q + 1.0);
        p0 = p1;
        p1 = p2;
    }
    *pn = p1;
    *dpn = n * (x * p1 - p0) / (x * x - 1.0);
}

static int gauss_rule(int n, double* a, double* w)
{
    if (n < 1 || n == 7 || n > 8) return -1;
    if (n == 1) { a[0] = 0.0; w[0] = 2.0; return 0; }
    const double pi = acos(-1.0);
    int m = n / 2;
    double r[4], rw[4], pn, dp;
    for (int k = 0; k < m; ++k) {
        double x = cos(pi * (k + 0.75) / (n + 0.5));  /* near the k-th largest root */
        for (int it = 0; it < 100; ++it) {
            legendre(n, x, &pn, &dp);
            double step = pn / dp;
            x -= step;
            if (fabs(step) < 1e-17) break;
        }
        legendre(n, x, &pn, &dp);
        r[m - 1 - k] = x;  /* increasing */
        rw[m - 1 - k] = 2.0 / ((1.0 - x * x) * dp * dp);
    }
    int q = 0;
    if (n & 1) {
        legendre(n, 0.0, &pn, &dp);
        a[q] = 0.0;
        w[q] = 2.0 / (dp * dp);
        ++q;
    }
    for (int k = 0; k < m; ++k) {
        if (n == 8) {
            a[q] = r[k]; w[q] = rw[k];
            a[q + m] = -r[k]; w[q + m] = rw[k];
            ++q;
        } else {
            a[q] = r[k]; w[q] = rw[k]; ++q;
            a[q] = -r[k]; w[q] = rw[k]; ++q;
        }
    }
    return 0;
}

int orc_gauss_rule(int n, double* a, double* w) { return gauss_rule(n, a, w); }

/* Quadratic MLS basis (GetBasis, pdDataManager2.cpp:1547-1562) at a point given relative
 * to the node and in units of dx.  The reference evaluates it at global coordinates; the
 * quadratic basis spans the same space after any shift and scaling, so the shape functions
 * are the same, and the moment matrix is far better conditioned (about 20 against 1e12). */
static void mls_basis(double x, double y, double z, double* b)
{
    b[0] = 1.0; b[1] = x; b[2] = y; b[3] = z;
    b[4] = x * y; b[5] = x * z; b[6] = y * z;
    b[7] = x * x; b[8] = y * y; b[9] = z * z;
}

/* Quartic spline weight (GetWeight, pdDataManager2.cpp:1564-1577) */
static double mls_weight(const double* xk, const double* xg, double rw)
{
    double a = xg[0] - xk[0], b = xg[1] - xk[1], c = xg[2] - xk[2];
    double r = sqrt(a * a + b * b + c * c);
    double d = r / rw;
    if (d > 1.0) return 0.0;
    return (1 - 6.0 * d * d + 8.0 * pow(d, 3.0) - 3.0 * pow(d, 4.0));
}

/* Gauss-Jordan elimination with full pivoting on the 10 x 10 system A x = b (the solver
 * of pdMatrix::Gauss, pdMatrix.cpp:251-340).  A and b are overwritten.  -1 if singular. */
static int gj_solve(double A[10][10], double* b, double* x)
{
    int colp[10];
    for (int k = 0; k < 10; ++k) colp[k] = k;
    for (int k = 0; k < 10; ++k) {
        int pr = k, pc = k;
        double big = -1.0;
        for (int r = k; r < 10; ++r)
            for (int c = k; c < 10; ++c)
                if (fabs(A[r][c]) > big) { big = fabs(A[r][c]); pr = r; pc = c; }
        if (!(big > 0.0)) return -1;
        if (pr != k) {
            for (int c = 0; c < 10; ++c) { double t = A[k][c]; A[k][c] = A[pr][c]; A[pr][c] = t; }
            double t = b[k]; b[k] = b[pr]; b[pr] = t;
        }
        if (pc != k) {
            for (int r = 0; r < 10; ++r) { double t = A[r][k]; A[r][k] = A[r][pc]; A[r][pc] = t; }
            int t = colp[k]; colp[k] = colp[pc]; colp[pc] = t;
        }
        double piv = A[k][k];
        for (int c = 0; c < 10; ++c) A[k][c] /= piv;
        b[k] /= piv;
        for (int r = 0; r < 10; ++r) {
            if (r == k || A[r][k] == 0.0) continue;
            double f = A[r][k];
            for (int c = 0; c < 10; ++c) A[r][c] -= f * A[k][c];
            b[r] -= f * b[k];
        }
    }
    for (int k = 0; k < 10; ++k) x[colp[k]] = b[k];
    return 0;
}

/* MLS contributing nodes of every node (InitIntegrationPoint, pdDataManager2.cpp:106-160):
 * the node itself, then its family within rw, or within 1.5 rw when that leaves fewer than
 * 10.  Returns the total list length. */
static int64_t init_contrib(orc_state* s)
{
    const orc_problem* p = &s->p;
    int64_t n = s->n;
    s->cb_rp = (int64_t*)malloc(sizeof(int64_t) * (n + 1));
    s->cb_rp[0] = 0;
    ivec lst = {0, 0, NULL};
    for (int64_t i = 0; i < n; ++i) {
        const double* xi = &s->X[3 * i];
        for (int pass = 0; pass < 2; ++pass) {
            double rad = pass ? p->rw * 1.5 : p->rw;
            int64_t start = lst.cnt;
            ivec_push(&lst, (int32_t)i);
            for (int64_t e = s->rp[i]; e < s->rp[i + 1]; ++e) {
                const double* xj = &s->X[3 * s->col[e]];
                double d = sqrt(sq(xj[0] - xi[0]) + sq(xj[1] - xi[1]) + sq(xj[2] - xi[2]));
                if (d <= rad) ivec_push(&lst, s->col[e]);
            }
            if (lst.cnt - start >= 10 || pass == 1) break;
            lst.cnt = start;
        }
        s->cb_rp[i + 1] = lst.cnt;
    }
    s->cb = lst.v;
    return lst.cnt;
}

/* MLS coefficients at the point xg of node o's contributing nodes, the basis relative to node
 * o in units of dx (InitMLSCoeff, pdIntgManager2.cpp:1879-1930; GetMLSCoeff's FGI twin,
 * pdDataManager2.cpp:177-254): phi_k = w_k p_k^T A^-1 p(xg), A = sum_k w_k p_k p_k^T.
 * -1 if A is singular. */
static int mls_coef_at(const orc_state* s, int64_t o, const double* xg, double* out)
{
    const orc_problem* p = &s->p;
    const double* xo = &s->X[3 * o];
    int64_t c0 = s->cb_rp[o], m = s->cb_rp[o + 1] - c0;
    double A[10][10], bb[10], sol[10], pg[10];
    memset(A, 0, sizeof A);
    for (int64_t k = 0; k < m; ++k) {
        const double* xk = &s->X[3 * s->cb[c0 + k]];
        double pk[10];
        mls_basis((xk[0] - xo[0]) / p->dx, (xk[1] - xo[1]) / p->dx, (xk[2] - xo[2]) / p->dx, pk);
        double wk = mls_weight(xk, xg, p->rw);
        for (int r = 0; r < 10; ++r)
            for (int c = 0; c < 10; ++c) A[r][c] += wk * pk[r] * pk[c];
    }
    mls_basis((xg[0] - xo[0]) / p->dx, (xg[1] - xo[1]) / p->dx, (xg[2] - xo[2]) / p->dx, pg);
    memcpy(bb, pg, sizeof bb);
    if (gj_solve(A, bb, sol)) return -1;
    for (int64_t k = 0; k < m; ++k) {
        const double* xk = &s->X[3 * s->cb[c0 + k]];
        double pk[10];
        mls_basis((xk[0] - xo[0]) / p->dx, (xk[1] - xo[1]) / p->dx, (xk[2] - xo[2]) / p->dx, pk);
        double v = 0.0;
        for (int r = 0; r < 10; ++r) v += pk[r] * sol[r];
        out[k] = mls_weight(xk, xg, p->rw) * v;
    }
    return 0;
}

/* SetGaussianPoint and InitIntegrationPoint's FGI branch (pdDataManager2.cpp:72-104,
 * 106-256): n^3 points per node (x index outermost, z innermost) at x_i + (dx/2) a, zero
 * displacement; for n > 1 the MLS contributing nodes (the node itself, then its family
 * within rw, or within 1.5 rw when that leaves fewer than 10) and their coefficients at
 * every point.  Every weight uses rw, as the reference's per-node support size does. */
static int init_fgi(orc_state* s)
{
    const orc_problem* p = &s->p;
    int n1 = p->intg_point;
    double a[8], w[8];
    if (gauss_rule(n1, a, w)) return -6;
    int ng = n1 * n1 * n1;
    int64_t n = s->n;
    s->ng = ng;
    s->gofs = (double*)malloc(sizeof(double) * 3 * ng);
    s->gw = (double*)malloc(sizeof(double) * 3 * ng);
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n1; ++j)
            for (int k = 0; k < n1; ++k) {
                int g = (i * n1 + j) * n1 + k;
                s->gofs[3 * g] = 0.5 * p->dx * a[i];
                s->gofs[3 * g + 1] = 0.5 * p->dx * a[j];
                s->gofs[3 * g + 2] = 0.5 * p->dx * a[k];
                s->gw[3 * g] = w[i]; s->gw[3 * g + 1] = w[j]; s->gw[3 * g + 2] = w[k];
            }
    s->G = (double*)calloc(3 * (size_t)(n ? n : 1) * ng, sizeof(double));
    if (n1 == 1) return 0;
    int64_t total = init_contrib(s);
    s->phi = (double*)malloc(sizeof(double) * (size_t)ng * (size_t)(total ? total : 1));
    int bad = 0;
    for (int64_t i = 0; i < n && !bad; ++i) {
        const double* xi = &s->X[3 * i];
        int64_t c0 = s->cb_rp[i], m = s->cb_rp[i + 1] - c0;
        double* ph = s->phi + (size_t)ng * c0;
        for (int g = 0; g < ng && !bad; ++g) {
            double xg[3] = {s->gofs[3 * g] + xi[0], s->gofs[3 * g + 1] + xi[1], s->gofs[3 * g + 2] + xi[2]};
            bad = mls_coef_at(s, i, xg, ph + (size_t)g * m) != 0;
        }
    }
    return bad ? -7 : 0;
}

/* CalcIntgPointDisplRK4 (pdDataManager.cpp:2402-2495) with UpdateIntgPointStatusWithBC
 * (:2604-2703).  mode 0: as the reference (BC nodes: DISPL / DISPLGRAD components take the
 * prescribed value at time t, other components keep theirs; other nodes: MLS interpolation of
 * u, or a copy for n = 1); mode 1: only the non-BC nodes; mode 2: every node interpolated. */
static int bc_node(const orc_state* s, int64_t i);
static void gauss_update(orc_state* s, double t, const double* u, int mode)
{
    const orc_problem* p = &s->p;
    int ng = s->ng;
    for (int64_t i = 0; i < s->n; ++i) {
        double* G = s->G + (size_t)3 * ng * i;
        if (mode != 2 && bc_node(s, i)) {
            if (mode == 1) continue;
            const orc_bc* bc = &p->bc[s->bc[i]];
            double fac = orc_time_factor(bc, t);
            for (int g = 0; g < ng; ++g) {
                double xg[3];
                for (int c = 0; c < 3; ++c) xg[c] = s->gofs[3 * g + c] + s->X[3 * i + c];
                for (int c = 0; c < 3; ++c) {
                    if (bc->f[c] != 1) continue;
                    if (bc->type == 1) G[3 * g + c] = bc->v[c] * fac;
                    else if (bc->type == 3)
                        G[3 * g + c] = (bc->v[3 * c] * xg[0] + bc->v[3 * c + 1] * xg[1] + bc->v[3 * c + 2] * xg[2]) * fac;
                }
            }
            continue;
        }
        if (!s->phi) {
            for (int c = 0; c < 3; ++c) G[c] = u[3 * i + c];
            continue;
        }
        int64_t c0 = s->cb_rp[i], m = s->cb_rp[i + 1] - c0;
        const double* ph = s->phi + (size_t)ng * c0;
        for (int g = 0; g < ng; ++g) {
            double a0 = 0.0, a1 = 0.0, a2 = 0.0;
            for (int64_t k = 0; k < m; ++k) {
                int64_t j = s->cb[c0 + k];
                double f = ph[(size_t)g * m + k];
                a0 += u[3 * j] * f;
                a1 += u[3 * j + 1] * f;
                a2 += u[3 * j + 2] * f;
            }
            G[3 * g] = a0; G[3 * g + 1] = a1; G[3 * g + 2] = a2;
        }
    }
}

/* The fixed-Gaussian branch of CalcBondBasedMotionRK4's family loop (pdDataManager.cpp:
 * 1918-1957) for the entry (i, j): the CCI bond force at each of node j's Gauss points with
 * that point's effective weight (w1 w2 w3)^(1/3), summed with the weights dx^3/8 w1 w2 w3. */
static void fgi_entry(const orc_state* s, int64_t i, int64_t j, const double* ui, double* F, double* el)
{
    const orc_problem* p = &s->p;
    const orc_mat* mi = &p->mat[s->mat[i]];
    const orc_mat* mj = &p->mat[s->mat[j]];
    double ecrit = (mj->ecrit < mi->ecrit) ? mj->ecrit : mi->ecrit;
    double ci = s->spring[s->mat[i]], cj = s->spring[s->mat[j]];
    double spring = (cj < ci) ? cj : ci;
    double bij = 0.5 * (s->be[i] + s->be[j]);
    const double* G = s->G + (size_t)3 * s->ng * j;
    F[0] = F[1] = F[2] = 0.0;
    for (int g = 0; g < s->ng; ++g) {
        double xg[3], xi[3], eta[3];
        for (int c = 0; c < 3; ++c) {
            xg[c] = s->gofs[3 * g + c] + s->X[3 * j + c];
            xi[c] = xg[c] - s->X[3 * i + c];
            eta[c] = G[3 * g + c] - ui[c];
        }
        double w1 = s->gw[3 * g], w2 = s->gw[3 * g + 1], w3 = s->gw[3 * g + 2];
        double effWeight = pow(w1 * w2 * w3, 1.0 / 3.0);
        double f[3], e1;
        orc_bond_force_cci(p, xi, eta, ecrit, spring, bij, effWeight, f, &e1);
        *el += e1;
        for (int c = 0; c < 3; ++c) F[c] += (pow(p->dx, 3) / 8.0) * w1 * w2 * w3 * f[c];
    }
}

/* ------------------------------------------------------------------ */
/* Adaptive integration (ai_on 1): CalcBondForceIntg (pdDataManager2.cpp:1222-1291) integrates
 * the bond force over node j's cell clipped to node i's horizon sphere with nested trapezoidal
 * rules (CalcIntg3D / CalcIntgX1 / CalcIntgX2, pdIntgManager.cpp:49-280; CalcTrapzdQuad,
 * :347-451): intg_point n gives 2^(n-1) + 1 points per direction, end points weighted 1/2.
 * The cell is taken whole when the bond's longest distance is inside the horizon
 * (CalcConfigurationType3, :3471-3482), and nothing when its shortest is not.  A partial cell
 * is integrated in x (outer), y, z (inner) between the cell faces and the sphere: the limits of
 * each direction are the face or the sphere, whichever is nearer, at the outer coordinates.
 * (The reference enumerates the sphere/cube configurations instead -- 5 subtypes with a node
 * on an axis, 39 cases otherwise, :523-3469 -- and returns 0 for a configuration outside its
 * list; here one rule covers every configuration.)  An interval shorter than 1e-10 adds
 * nothing, as in the reference.  A point's weight is the product of its three 1-D weights. */
static int ai_cell_points(const orc_problem* p, const double* c, double dmin, double dmax, double* xw, int cap)
{
    double delta = p->horizon, h = 0.5 * p->dx;
    if (delta <= dmin) return 0;
    int full = delta >= dmax;
    int nseg = 1 << (p->intg_point - 1);
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) { lo[k] = c[k] - h; hi[k] = c[k] + h; }
    double a1 = lo[0], b1 = hi[0];
    if (!full) { a1 = fmax(a1, -delta); b1 = fmin(b1, delta); }
    if (b1 - a1 <= 1.0e-10) return 0;
    double H1 = (b1 - a1) / nseg;
    int cnt = 0;
    for (int ia = 0; ia <= nseg; ++ia) {
        double x = ((nseg - ia) * a1 + ia * b1) / nseg;
        double w1 = H1 * ((ia == 0 || ia == nseg) ? 0.5 : 1.0);
        double a2 = lo[1], b2 = hi[1];
        if (!full) {
            double r1 = sqrt(fmax(delta * delta - x * x, 0.0));
            a2 = fmax(a2, -r1); b2 = fmin(b2, r1);
        }
        if (b2 - a2 <= 1.0e-10) continue;
        double H2 = (b2 - a2) / nseg;
        for (int ib = 0; ib <= nseg; ++ib) {
            double y = ((nseg - ib) * a2 + ib * b2) / nseg;
            double w2 = H2 * ((ib == 0 || ib == nseg) ? 0.5 : 1.0);
            double a3 = lo[2], b3 = hi[2];
            if (!full) {
                double r2 = sqrt(fmax(delta * delta - x * x - y * y, 0.0));
                a3 = fmax(a3, -r2); b3 = fmin(b3, r2);
            }
            if (b3 - a3 <= 1.0e-10) continue;
            double H3 = (b3 - a3) / nseg;
            for (int ic = 0; ic <= nseg; ++ic) {
                double z = ((nseg - ic) * a3 + ic * b3) / nseg;
                double w3 = H3 * ((ic == 0 || ic == nseg) ? 0.5 : 1.0);
                if (cnt < cap) {
                    xw[4 * cnt] = x; xw[4 * cnt + 1] = y; xw[4 * cnt + 2] = z;
                    xw[4 * cnt + 3] = w1 * w2 * w3;
                }
                ++cnt;
            }
        }
    }
    return cnt;
}

int orc_ai_cell_points(const orc_problem* p, const double* c, double dmin, double dmax, double* xw, int cap)
{
    return ai_cell_points(p, c, dmin, dmax, xw, cap);
}

/* The bond's shortest and longest distances as InitBond stores them (pdDataManager2.cpp:33-53):
 * from the lower-id node to the higher-id node's cell, shared by both directions. */
static void bond_dists(const orc_state* s, int64_t i, int64_t j, double* dmin, double* dmax)
{
    int64_t a = (i < j) ? i : j, b = (i < j) ? j : i;
    dist_to_cube(&s->X[3 * a], &s->X[3 * b], s->p.dx, dmin, dmax);
}

/* Points of every uncut entry (reference CSR order), then with mls_on the MLS coefficients of
 * each point over node j's contributing nodes (InitIntegrationPoint's AI branch,
 * pdDataManager2.cpp:257-297, which runs the same integration with a unit integrand to place
 * them).  Without MLS every node must carry a DISPLGRAD condition: the integrand prescribes the
 * points' relative displacement from it and exits otherwise (pdIntgManager2.cpp:187-211). */
static int init_ai(orc_state* s)
{
    const orc_problem* p = &s->p;
    int64_t n = s->n, E = s->rp[n];
    if (p->intg_point < 1 || p->intg_point > 5) return -8;
    if (!p->mls_on)
        for (int64_t i = 0; i < n; ++i)
            if (s->bc[i] < 0 || p->bc[s->bc[i]].type != 3) return -9;
    int nmax = (1 << (p->intg_point - 1)) + 1;
    nmax = nmax * nmax * nmax;
    s->ai = 1;
    s->ai_ep = (int64_t*)malloc(sizeof(int64_t) * (E + 1));
    s->ai_ep[0] = 0;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t e = s->rp[i]; e < s->rp[i + 1]; ++e) {
            int cnt = 0;
            if (!s->cut[e]) {
                int64_t j = s->col[e];
                double dmin, dmax, c[3];
                bond_dists(s, i, j, &dmin, &dmax);
                for (int k = 0; k < 3; ++k) c[k] = s->X[3 * j + k] - s->X[3 * i + k];
                cnt = ai_cell_points(p, c, dmin, dmax, NULL, 0);
            }
            s->ai_ep[e + 1] = s->ai_ep[e] + cnt;
        }
    int64_t P = s->ai_ep[E];
    s->ai_xw = (double*)malloc(sizeof(double) * 4 * (size_t)(P ? P : 1));
    if (!s->ai_xw) return -10;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t e = s->rp[i]; e < s->rp[i + 1]; ++e) {
            if (s->ai_ep[e + 1] == s->ai_ep[e]) continue;
            int64_t j = s->col[e];
            double dmin, dmax, c[3];
            bond_dists(s, i, j, &dmin, &dmax);
            for (int k = 0; k < 3; ++k) c[k] = s->X[3 * j + k] - s->X[3 * i + k];
            ai_cell_points(p, c, dmin, dmax, s->ai_xw + 4 * s->ai_ep[e], nmax);
        }
    if (!p->mls_on) return 0;
    init_contrib(s);
    s->ai_po = (int64_t*)malloc(sizeof(int64_t) * (P + 1));
    s->ai_po[0] = 0;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t e = s->rp[i]; e < s->rp[i + 1]; ++e) {
            int64_t j = s->col[e];
            int64_t m = s->cb_rp[j + 1] - s->cb_rp[j];
            for (int64_t q = s->ai_ep[e]; q < s->ai_ep[e + 1]; ++q) s->ai_po[q + 1] = s->ai_po[q] + m;
        }
    s->ai_phi = (double*)malloc(sizeof(double) * (size_t)(s->ai_po[P] ? s->ai_po[P] : 1));
    if (!s->ai_phi) return -10;
    int bad = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4) reduction(| : bad)
#endif
    for (int64_t i = 0; i < n; ++i)
        for (int64_t e = s->rp[i]; e < s->rp[i + 1] && !bad; ++e) {
            int64_t j = s->col[e];
            for (int64_t q = s->ai_ep[e]; q < s->ai_ep[e + 1] && !bad; ++q) {
                /* the point in global coordinates (InitMLSCoeff, pdIntgManager2.cpp:1891-1893) */
                double xg[3];
                for (int k = 0; k < 3; ++k) xg[k] = s->ai_xw[4 * q + k] + s->X[3 * i + k];
                bad = mls_coef_at(s, j, xg, s->ai_phi + s->ai_po[q]) != 0;
            }
        }
    return bad ? -7 : 0;
}

/* CalcBondBasedForce (pdIntgManager2.cpp:262-309) at one point xw (x, y, z relative to node i,
 * weight) with beta = 1, added with the point's weight; energy scaled by weight / vol. */
static void ai_point_force(const double* xw, const double* eta, double ecrit, double spring, double bij, double vol,
                           double* F, double* el)
{
    double xi1 = xw[0], xi2 = xw[1], xi3 = xw[2];
    double r = sqrt(xi1 * xi1 + xi2 * xi2 + xi3 * xi3);
    double pp = sqrt(sq(xi1 + eta[0]) + sq(xi2 + eta[1]) + sq(xi3 + eta[2]));
    double u = pp - r;
    double stretch = u / r;
    if (u <= 0.0 || (u > 0.0 && stretch <= ecrit)) {
        double a = bij * 1.0 * spring * stretch;
        F[0] += xw[3] * (a * (xi1 + eta[0]) / pp);
        F[1] += xw[3] * (a * (xi2 + eta[1]) / pp);
        F[2] += xw[3] * (a * (xi3 + eta[2]) / pp);
        *el += bij * 0.5 * (0.5 * 1.0 * (spring / r) * u * u) * (xw[3] / vol);
    }
}

/* The AI branch of CalcBondBasedMotionRK4's family loop (pdDataManager.cpp:1888-1916) for entry
 * e = (i, j): CalcBondBasedForce (pdIntgManager2.cpp:92-310) at every point, beta = 1, summed
 * with the points' weights.  The point's displacement: with MLS, a BC node j's DISPL / DISPLGRAD
 * value at time t in its flagged components (0 in the others; mode 2 skips this), else the MLS
 * interpolation of the contributing nodes' trial displacements `trial`; without MLS the
 * relative displacement is G_j xi (node j's DISPLGRAD gradient).  The elastic energy density
 * of a point is scaled by its share of a cell, weight / dx^3 (the reference adds it unscaled,
 * pdIntgManager2.cpp:287-292). */
static void ai_entry(const orc_state* s, int64_t i, int64_t e, const double* ui, const double* trial, double t,
                     int mode, double* F, double* el)
{
    const orc_problem* p = &s->p;
    int64_t j = s->col[e];
    const orc_mat* mi = &p->mat[s->mat[i]];
    const orc_mat* mj = &p->mat[s->mat[j]];
    double ecrit = (mj->ecrit < mi->ecrit) ? mj->ecrit : mi->ecrit;
    double ci = s->spring[s->mat[i]], cj = s->spring[s->mat[j]];
    double spring = (cj < ci) ? cj : ci;
    double bij = 0.5 * (s->be[i] + s->be[j]);
    double vol = p->dx * p->dx * p->dx;
    const orc_bc* bcj = (s->bc[j] >= 0) ? &p->bc[s->bc[j]] : NULL;
    int prescribed = p->mls_on && mode != 2 && bc_node(s, j);
    double fac = prescribed ? orc_time_factor(bcj, t) : 0.0;
    F[0] = F[1] = F[2] = 0.0;
    for (int64_t q = s->ai_ep[e]; q < s->ai_ep[e + 1]; ++q) {
        const double* xw = &s->ai_xw[4 * q];
        double xi1 = xw[0], xi2 = xw[1], xi3 = xw[2];
        double eta[3];
        if (!p->mls_on) {
            const double* G = bcj->v;
            eta[0] = G[0] * xi1 + G[1] * xi2 + G[2] * xi3;
            eta[1] = G[3] * xi1 + G[4] * xi2 + G[5] * xi3;
            eta[2] = G[6] * xi1 + G[7] * xi2 + G[8] * xi3;
        } else {
            double uj[3] = {0.0, 0.0, 0.0};
            if (prescribed) {
                double xg[3] = {xi1 + s->X[3 * i], xi2 + s->X[3 * i + 1], xi3 + s->X[3 * i + 2]};
                for (int c = 0; c < 3; ++c) {
                    if (bcj->f[c] != 1) continue;
                    if (bcj->type == 1) uj[c] = bcj->v[c] * fac;
                    else if (bcj->type == 3)
                        uj[c] = (bcj->v[3 * c] * xg[0] + bcj->v[3 * c + 1] * xg[1] + bcj->v[3 * c + 2] * xg[2]) * fac;
                }
            } else {
                int64_t c0 = s->cb_rp[j], m = s->cb_rp[j + 1] - c0;
                const double* ph = s->ai_phi + s->ai_po[q];
                for (int64_t k = 0; k < m; ++k) {
                    int64_t o = s->cb[c0 + k];
                    uj[0] += trial[3 * o] * ph[k];
                    uj[1] += trial[3 * o + 1] * ph[k];
                    uj[2] += trial[3 * o + 2] * ph[k];
                }
            }
            for (int c = 0; c < 3; ++c) eta[c] = uj[c] - ui[c];
        }
        ai_point_force(xw, eta, ecrit, spring, bij, vol, F, el);
    }
}

int orc_ai_cell_force(const orc_problem* p, const double* c, double dmin, double dmax, const double* G, double bij,
                      double spring, double ecrit, double* F, double* el)
{
    int cnt = ai_cell_points(p, c, dmin, dmax, NULL, 0);
    double* xw = (double*)malloc(sizeof(double) * 4 * (size_t)(cnt ? cnt : 1));
    ai_cell_points(p, c, dmin, dmax, xw, cnt);
    double vol = p->dx * p->dx * p->dx;
    F[0] = F[1] = F[2] = 0.0;
    *el = 0.0;
    for (int q = 0; q < cnt; ++q) {
        const double* x = &xw[4 * q];
        double eta[3] = {G[0] * x[0] + G[1] * x[1] + G[2] * x[2], G[3] * x[0] + G[4] * x[1] + G[5] * x[2],
                         G[6] * x[0] + G[7] * x[1] + G[8] * x[2]};
        ai_point_force(x, eta, ecrit, spring, bij, vol, F, el);
    }
    free(xw);
    return cnt;
}

int orc_init(orc_state* s, int method, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    int rc = init_nodes(s);
    if (rc) return rc;
    rc = init_bonds(s, method, nthreads);
    if (rc) return rc;
    init_material(s);
    init_bdryeff(s, nthreads);
    init_traction(s);
    if (s->p.fgi_on && !s->p.cci_on) {
        rc = init_fgi(s);
        if (rc) return rc;
    } else if (s->p.ai_on && !s->p.cci_on) {
        rc = init_ai(s);
        if (rc) return rc;
    }
    int64_t n = s->n;
    s->U = (double*)calloc(3 * (n ? n : 1), sizeof(double));
    s->V = (double*)calloc(3 * (n ? n : 1), sizeof(double));
    s->Vb = (double*)calloc(3 * (n ? n : 1), sizeof(double));
    s->elast = (double*)calloc((n ? n : 1), sizeof(double));
    s->dt_stable = s->p.fixed_dt_on ? s->p.fixed_dt : s->p.safety * s->min_bond / s->max_c;
    s->dt_cur = s->dt_stable;
    s->t = 0.0;
    s->steps_done = 0;
    s->last_step = s->p.total_step;
    return 0;
}

int64_t orc_num_nodes(const orc_state* s) { return s->n; }
int64_t orc_num_entries(const orc_state* s) { return s->rp ? s->rp[s->n] : 0; }
int orc_num_max_family(const orc_state* s) { return s->max_family; }
double orc_min_bond_length(const orc_state* s) { return s->min_bond; }
double orc_stable_dt(const orc_state* s) { return s->dt_stable; }
void orc_get_nodes(const orc_state* s, double* X, int* mat, int* bc)
{
    if (X) memcpy(X, s->X, sizeof(double) * 3 * s->n);
    if (mat) memcpy(mat, s->mat, sizeof(int) * s->n);
    if (bc) memcpy(bc, s->bc, sizeof(int) * s->n);
}
void orc_get_csr(const orc_state* s, int64_t* rp, int32_t* col, uint8_t* cut)
{
    if (rp) memcpy(rp, s->rp, sizeof(int64_t) * (s->n + 1));
    if (col) memcpy(col, s->col, sizeof(int32_t) * s->rp[s->n]);
    if (cut) memcpy(cut, s->cut, s->rp[s->n]);
}
void orc_get_bdryeff(const orc_state* s, double* be) { memcpy(be, s->be, sizeof(double) * s->n); }
void orc_set_bdryeff(orc_state* s, const double* be) { memcpy(s->be, be, sizeof(double) * s->n); }

/* The CCI branch of the family loop of CalcBondBasedMotionRK4 (pdDataManager.cpp:1795-1864),
 * for node i with the displacement of every node read from u. */
static void row_force(const orc_state* s, int64_t i, const int32_t* fam, const uint8_t* fcut, int64_t cnt,
                      const double* u, double* F, double* el)
{
    const orc_problem* p = &s->p;
    const orc_mat* mi = &p->mat[s->mat[i]];
    double acc[3] = {0.0, 0.0, 0.0}, ea = 0.0;
    for (int64_t e = 0; e < cnt; ++e) {
        if (fcut[e]) continue;
        int64_t j = fam[e];
        const orc_mat* mj = &p->mat[s->mat[j]];
        double xi[3] = {s->X[3 * j] - s->X[3 * i], s->X[3 * j + 1] - s->X[3 * i + 1], s->X[3 * j + 2] - s->X[3 * i + 2]};
        double eta[3] = {u[3 * j] - u[3 * i], u[3 * j + 1] - u[3 * i + 1], u[3 * j + 2] - u[3 * i + 2]};
        double ecrit = (mj->ecrit < mi->ecrit) ? mj->ecrit : mi->ecrit;
        double ci = s->spring[s->mat[i]], cj = s->spring[s->mat[j]];
        double spring = (cj < ci) ? cj : ci;
        double bij = 0.5 * (s->be[i] + s->be[j]);
        double f[3], el1;
        orc_bond_force_cci(p, xi, eta, ecrit, spring, bij, 2.0, f, &el1);
        acc[0] += f[0]; acc[1] += f[1]; acc[2] += f[2];
        ea += el1;
    }
    F[0] = acc[0]; F[1] = acc[1]; F[2] = acc[2];
    if (el) *el = ea;
}

static void family_force(const orc_state* s, int64_t i, const double* u, double* F, double* el)
{
    row_force(s, i, &s->col[s->rp[i]], &s->cut[s->rp[i]], s->rp[i + 1] - s->rp[i], u, F, el);
}

void orc_forces(const orc_state* s, const double* u, double* F, double* elast, int nthreads)
{
    if (s->ai) {
        /* AI: every node's points interpolated from u (no BC values) */
#ifdef _OPENMP
        if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16) if (nthreads != 1)
#endif
        for (int64_t i = 0; i < s->n; ++i) {
            double acc[3] = {0.0, 0.0, 0.0}, ea = 0.0;
            for (int64_t e = s->rp[i]; e < s->rp[i + 1]; ++e) {
                if (s->cut[e]) continue;
                double f[3];
                ai_entry(s, i, e, &u[3 * i], u, 0.0, 2, f, &ea);
                acc[0] += f[0]; acc[1] += f[1]; acc[2] += f[2];
            }
            for (int c = 0; c < 3; ++c) F[3 * i + c] = acc[c];
            if (elast) elast[i] = ea;
        }
        return;
    }
    if (s->ng) {
        /* FGI: every node's Gauss points interpolated from u (no BC values), in a scratch
         * copy so that the stepping state keeps its own */
        orc_state* w = (orc_state*)s;
        double* keep = w->G;
        w->G = (double*)malloc(sizeof(double) * 3 * (size_t)(s->n ? s->n : 1) * s->ng);
        gauss_update(w, 0.0, u, 2);
        for (int64_t i = 0; i < s->n; ++i) {
            double acc[3] = {0.0, 0.0, 0.0}, ea = 0.0;
            for (int64_t e = s->rp[i]; e < s->rp[i + 1]; ++e) {
                if (s->cut[e]) continue;
                double f[3];
                fgi_entry(s, i, s->col[e], &u[3 * i], f, &ea);
                acc[0] += f[0]; acc[1] += f[1]; acc[2] += f[2];
            }
            for (int c = 0; c < 3; ++c) F[3 * i + c] = acc[c];
            if (elast) elast[i] = ea;
        }
        free(w->G);
        w->G = keep;
        return;
    }
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256) if (nthreads != 1)
#endif
    for (int64_t i = 0; i < s->n; ++i) family_force(s, i, u, &F[3 * i], elast ? &elast[i] : NULL);
}

/* Tangent (eta -> 0) of the same CCI bond force: (K u)_i = sum_j w_ij xi xi^T (u_i - u_j) / r^3
 * with w_ij = bdryEff_ij * dx^3 * beta * c_ij, i.e. the operator of CalcDampingCoeff
 * (pdDataManager2.cpp:1763-1852) with the force's beta and bdryEff weights. */
void orc_apply_tangent(const orc_state* s, const double* u, double* y, int nthreads)
{
    const orc_problem* p = &s->p;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256) if (nthreads != 1)
#endif
    for (int64_t i = 0; i < s->n; ++i) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (int64_t e = s->rp[i]; e < s->rp[i + 1]; ++e) {
            if (s->cut[e]) continue;
            int64_t j = s->col[e];
            double xi[3] = {s->X[3 * j] - s->X[3 * i], s->X[3 * j + 1] - s->X[3 * i + 1], s->X[3 * j + 2] - s->X[3 * i + 2]};
            double r = sqrt(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]);
            double ci = s->spring[s->mat[i]], cj = s->spring[s->mat[j]];
            double c = (cj < ci) ? cj : ci;
            double bij = 0.5 * (s->be[i] + s->be[j]);
            double beta = orc_beta(r, p->horizon, p->dx, p->cci_on, p->cci_mdf);
            double w = bij * (p->dx * p->dx * p->dx) * beta * c;
            double d[3] = {u[3 * i] - u[3 * j], u[3 * i + 1] - u[3 * j + 1], u[3 * i + 2] - u[3 * j + 2]};
            double proj = (xi[0] * d[0] + xi[1] * d[1] + xi[2] * d[2]) * w / (r * r * r);
            acc[0] += proj * xi[0]; acc[1] += proj * xi[1]; acc[2] += proj * xi[2];
        }
        y[3 * i] = acc[0]; y[3 * i + 1] = acc[1]; y[3 * i + 2] = acc[2];
    }
}

/* A node is a "BC node" (pdNode::GetBC() non-null) when it lies in a boundary space.  With the
 * traction-as-body-force extension on, TRACTION nodes stay free nodes: in the reference a type-4
 * node keeps its stale U through all four stages with no overwrite, an update whose amplification
 * has determinant 1 + 0.875 (w dt)^2 > 1, i.e. unconditionally unstable. */
static int bc_node(const orc_state* s, int64_t i)
{
    int b = s->bc[i];
    if (b < 0) return 0;
    if (s->p.traction_body && s->p.bc[b].type == 4) return 0;
    return 1;
}

/* UpdateNodeStatusWithBC, pdDataManager.cpp:2497-2602 */
static void apply_bc(orc_state* s, int64_t i, double t)
{
    int b = s->bc[i];
    if (b < 0) return;
    const orc_bc* bc = &s->p.bc[b];
    double fac = orc_time_factor(bc, t);
    double* U = &s->U[3 * i];
    double* Vb = &s->Vb[3 * i];
    const double* X = &s->X[3 * i];
    for (int c = 0; c < 3; ++c) {
        if (bc->f[c] != 1) continue;
        if (bc->type == 1) { U[c] = bc->v[c] * fac; Vb[c] = 0.0; }
        else if (bc->type == 2) { Vb[c] = bc->v[c] * fac; }
        else if (bc->type == 3) {
            U[c] = (bc->v[3 * c] * X[0] + bc->v[3 * c + 1] * X[1] + bc->v[3 * c + 2] * X[2]) * fac;
            Vb[c] = 0.0;
        }
    }
}

/* One stage of CalcBondBasedMotionRK4 (pdDataManager.cpp:1767-1992): Vbuf accumulation per bond,
 * BC overwrite, k = Vbuf * dt_s.  trial: displStep (3N). */
static void rk4_stage(orc_state* s, double dts, double t, const double* trial, double* k, int nthreads)
{
    const orc_problem* p = &s->p;
    int64_t n = s->n;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256) if (nthreads != 1)
#endif
    for (int64_t i = 0; i < n; ++i) {
        const orc_mat* mi = &p->mat[s->mat[i]];
        double denst = mi->rho;
        double bf[3] = {mi->g[0] * denst, mi->g[1] * denst, mi->g[2] * denst};
        const double* ui = bc_node(s, i) ? &s->U[3 * i] : &trial[3 * i];
        for (int64_t e = s->rp[i]; e < s->rp[i + 1]; ++e) {
            if (s->cut[e]) continue;
            int64_t j = s->col[e];
            if (s->ai) {
                /* adaptive integration: the entry's points, displacements from this stage's trial */
                double f[3], el = 0.0;
                ai_entry(s, i, e, ui, trial, t, 0, f, &el);
                s->elast[i] += el;
                for (int c = 0; c < 3; ++c) s->Vb[3 * i + c] += (f[c] + bf[c]) * dts / denst;
                continue;
            }
            if (s->ng) {
                /* fixed Gaussian integration: node j's Gauss points */
                double f[3], el = 0.0;
                fgi_entry(s, i, j, ui, f, &el);
                s->elast[i] += el;
                for (int c = 0; c < 3; ++c) s->Vb[3 * i + c] += (f[c] + bf[c]) * dts / denst;
                continue;
            }
            const orc_mat* mj = &p->mat[s->mat[j]];
            double xi[3] = {s->X[3 * j] - s->X[3 * i], s->X[3 * j + 1] - s->X[3 * i + 1], s->X[3 * j + 2] - s->X[3 * i + 2]};
            const double* uj = bc_node(s, j) ? &s->U[3 * j] : &trial[3 * j];
            double eta[3] = {uj[0] - ui[0], uj[1] - ui[1], uj[2] - ui[2]};
            double ecrit = (mj->ecrit < mi->ecrit) ? mj->ecrit : mi->ecrit;
            double ci = s->spring[s->mat[i]], cj = s->spring[s->mat[j]];
            double spring = (cj < ci) ? cj : ci;
            double bij = 0.5 * (s->be[i] + s->be[j]);
            double f[3], el;
            orc_bond_force_cci(p, xi, eta, ecrit, spring, bij, 2.0, f, &el);
            s->elast[i] += el;
            s->Vb[3 * i] += (f[0] + bf[0]) * dts / denst;
            s->Vb[3 * i + 1] += (f[1] + bf[1]) * dts / denst;
            s->Vb[3 * i + 2] += (f[2] + bf[2]) * dts / denst;
        }
        /* traction body force (extension), once per node */
        if (p->traction_body && s->bc[i] >= 0 && p->bc[s->bc[i]].type == 4) {
            double fac = orc_time_factor(&p->bc[s->bc[i]], t);
            for (int c = 0; c < 3; ++c) s->Vb[3 * i + c] += s->trac[3 * i + c] * fac * dts / denst;
        }
    }
    /* the BC overwrite and k for all nodes happen after the whole family loop */
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (nthreads != 1)
#endif
    for (int64_t i = 0; i < n; ++i) {
        apply_bc(s, i, t);
        k[3 * i] = s->Vb[3 * i] * dts;
        k[3 * i + 1] = s->Vb[3 * i + 1] * dts;
        k[3 * i + 2] = s->Vb[3 * i + 2] * dts;
    }
}

/* RK4, pdDataManager.cpp:1994-2134.  CalcIntgPointDisplRK4 runs before stages 2-4 and after
 * the update (a no-op for CCI): stage 1 uses the Gauss-point displacements of the previous step. */
static void rk4(orc_state* s, double dt, double t, int nthreads)
{
    const orc_problem* p = &s->p;
    int64_t n = s->n, n3 = 3 * n;
    double* u1 = (double*)malloc(sizeof(double) * n3);
    double* u234 = (double*)malloc(sizeof(double) * n3);
    double* k1 = (double*)malloc(sizeof(double) * n3);
    double* k24 = (double*)malloc(sizeof(double) * n3);
    double* k3 = (double*)malloc(sizeof(double) * n3);
    double dtHalf = 0.5 * dt, dt6 = 1.0 / 6.0;
    memcpy(u1, s->U, sizeof(double) * n3);
    rk4_stage(s, dtHalf, t, u1, k1, nthreads);
    for (int64_t i = 0; i < n3; ++i) u234[i] = u1[i] + 0.5 * k1[i];
    if (s->ng) gauss_update(s, t, u234, 0);
    rk4_stage(s, dtHalf, t, u234, k24, nthreads);
    for (int64_t i = 0; i < n3; ++i) u234[i] = u1[i] + 0.5 * k24[i];
    if (s->ng) gauss_update(s, t, u234, 0);
    rk4_stage(s, dt, t, u234, k3, nthreads);
    for (int64_t i = 0; i < n3; ++i) { u234[i] = u1[i] + k3[i]; k3[i] += k24[i]; }
    if (s->ng) gauss_update(s, t, u234, 0);
    rk4_stage(s, dt, t, u234, k24, nthreads);
    for (int64_t i = 0; i < n; ++i) {
        double denst = p->mat[s->mat[i]].rho;
        for (int c = 0; c < 3; ++c) {
            int64_t r = 3 * i + c;
            if (bc_node(s, i))
                s->U[r] = s->U[r] + dt6 * (k1[r] + k24[r] + 2.0 * k3[r]);
            else
                s->U[r] = s->U[r] + dt6 * (k1[r] + k24[r] + 2.0 * k3[r]) - dt * p->damping * s->V[r] * dt / denst;
        }
        for (int c = 0; c < 3; ++c) s->Vb[3 * i + c] += -p->damping * s->V[3 * i + c] * dt / denst;
        for (int c = 0; c < 3; ++c) s->V[3 * i + c] = s->Vb[3 * i + c];
    }
    double vol = pow(p->dx, 3.0);
    for (int64_t i = 0; i < n; ++i) {
        double v1 = s->V[3 * i], v2 = s->V[3 * i + 1], v3 = s->V[3 * i + 2];
        double velsq = v1 * v1 + v2 * v2 + v3 * v3;
        double denst = p->mat[s->mat[i]].rho;
        s->ee += vol * s->elast[i];
        s->ke += 0.5 * denst * velsq * vol;
    }
    if (s->ng) gauss_update(s, t, s->U, 0);
    free(u1); free(u234); free(k1); free(k24); free(k3);
}

/* pdSolve::SolveDynamic loop body (pdSolve.cpp:55-120), nsteps iterations from the current one */
int orc_run_dynamic(orc_state* s, int nsteps, int nthreads)
{
    const orc_problem* p = &s->p;
    for (int it = 0; it < nsteps; ++it) {
        int iter = s->steps_done + 1;
        if (iter > p->total_step) return 1;
        s->t += s->dt_cur;
        if (!(s->t <= p->total_time)) { s->t -= s->dt_cur; return 1; }
        /* InitDynamicSolve, pdDataManager2.cpp:410-439 */
        s->ee = 0.0; s->ke = 0.0;
        memset(s->elast, 0, sizeof(double) * s->n);
        memcpy(s->Vb, s->V, sizeof(double) * 3 * s->n);
        rk4(s, s->dt_cur, s->t, nthreads);
        s->steps_done = iter;
        if (s->t + s->dt_stable > p->total_time && s->t != p->total_time) {
            s->last_step = iter + 1;
            s->dt_cur = p->total_time - s->t;
        }
    }
    return 0;
}

/* Static system of the build's static solve (the intent of pdSolve::AssembleEquation,
 * pdSolve.cpp:195-363, on the tangent of the CCI force): per component a free mask
 * (1 free, 0 prescribed), prescribed values uc and loads b.  DISPL flagged components
 * take value * factor(T), VELOC value * T, DISPLGRAD (G X) * factor(T), T = total_time;
 * b = body force once per unbroken bond + traction body force (extension). */
void orc_static_system(const orc_state* s, double* mask, double* uc, double* b)
{
    const orc_problem* p = &s->p;
    double T = p->total_time;
    for (int64_t i = 0; i < s->n; ++i) {
        int64_t nb = 0;
        for (int64_t e = s->rp[i]; e < s->rp[i + 1]; ++e) nb += s->cut[e] ? 0 : 1;
        const orc_mat* m = &p->mat[s->mat[i]];
        for (int c = 0; c < 3; ++c) {
            mask[3 * i + c] = 1.0;
            uc[3 * i + c] = 0.0;
            b[3 * i + c] = (double)nb * (m->g[c] * m->rho);
        }
        int bi = s->bc[i];
        if (bi < 0) continue;
        const orc_bc* bc = &p->bc[bi];
        double fac = orc_time_factor(bc, T);
        for (int c = 0; c < 3; ++c) {
            if (bc->type == 4) b[3 * i + c] += s->trac[3 * i + c] * fac;
            if (bc->f[c] != 1 || !bc_node(s, i)) continue;
            if (bc->type == 1) { mask[3 * i + c] = 0.0; uc[3 * i + c] = bc->v[c] * fac; }
            else if (bc->type == 2) { mask[3 * i + c] = 0.0; uc[3 * i + c] = bc->v[c] * T; }
            else if (bc->type == 3) {
                mask[3 * i + c] = 0.0;
                uc[3 * i + c] = (bc->v[3 * c] * s->X[3 * i] + bc->v[3 * c + 1] * s->X[3 * i + 1] +
                                 bc->v[3 * c + 2] * s->X[3 * i + 2]) * fac;
            }
        }
    }
}

/* Families, cut flags and boundary-effect factors of selected rows only, without building
 * the whole CSR: InitNode, then for each row the cell-list family search of InitBond
 * (pdDataManager2.cpp:12-70), the slit test and InitBdryEffectFactor's per-node factor
 * (pdDataManager.cpp:1340-1473).  Rows are written in the given order: rp (nrows + 1),
 * col / cut (cap entries), be (nrows).  Returns the entry count, -1 if cap is too small,
 * or a negative init code. */
int64_t orc_sample_rows(const orc_problem* p, const int64_t* rows, int64_t nrows, int64_t* rp, int32_t* col,
                        uint8_t* cut, double* be, int64_t cap, int nthreads)
{
    orc_state* s = orc_create(p);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    int rc = init_nodes(s);
    if (rc) {
        orc_destroy(s);
        return rc;
    }
    init_material(s);
    const int maxf = virtual_max_family(s);
    cell_grid g;
    grid_build(s, &g);
    ivec* fam = (ivec*)calloc(nrows ? nrows : 1, sizeof(ivec));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) if (nthreads != 1)
#endif
    for (int64_t k = 0; k < nrows; ++k) {
        int64_t i = rows[k];
        if (i < 0 || i >= s->n) continue;
        gather_family(s, &g, i, &fam[k], NULL);
        be[k] = p->bdryeff_on ? bdryeff_node(s, i, fam[k].v, fam[k].cnt, maxf) : 1.0;
    }
    grid_free(&g);
    int64_t tot = 0;
    rp[0] = 0;
    for (int64_t k = 0; k < nrows; ++k) {
        rp[k + 1] = rp[k] + fam[k].cnt;
        tot += fam[k].cnt;
    }
    if (tot <= cap) {
        for (int64_t k = 0; k < nrows; ++k) {
            int64_t i = rows[k];
            for (int64_t e = 0; e < fam[k].cnt; ++e) {
                int64_t j = fam[k].v[e];
                int64_t a = i < j ? i : j, b = i < j ? j : i;
                col[rp[k] + e] = (int32_t)j;
                cut[rp[k] + e] = (uint8_t)orc_pair_cut(p, &s->X[3 * a], &s->X[3 * b]);
            }
        }
    }
    for (int64_t k = 0; k < nrows; ++k) free(fam[k].v);
    free(fam);
    orc_destroy(s);
    return tot <= cap ? tot : -1;
}

/* Family force sums (CCI, no body force) of the given rows only, for a displacement field u
 * of every node (3N): the family loop of CalcBondBasedMotionRK4 (pdDataManager.cpp:1795-1864)
 * over each row's family from the cell-list search, with the bdryEff factors of the row and of
 * its neighbours computed from their own families.  F: 3 nrows, el: nrows (may be NULL). */
int orc_sample_forces(const orc_problem* p, const int64_t* rows, int64_t nrows, const double* u, double* F,
                      double* el, int nthreads)
{
    orc_state* s = orc_create(p);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    int rc = init_nodes(s);
    if (rc) {
        orc_destroy(s);
        return rc;
    }
    init_material(s);
    const int maxf = virtual_max_family(s);
    cell_grid g;
    grid_build(s, &g);
    ivec* fam = (ivec*)calloc(nrows ? nrows : 1, sizeof(ivec));
    uint8_t** fcut = (uint8_t**)calloc(nrows ? nrows : 1, sizeof(uint8_t*));
    s->be = (double*)malloc(sizeof(double) * (s->n ? s->n : 1));
    unsigned char* need = (unsigned char*)calloc(s->n ? s->n : 1, 1);
    for (int64_t i = 0; i < s->n; ++i) s->be[i] = 1.0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) if (nthreads != 1)
#endif
    for (int64_t k = 0; k < nrows; ++k) {
        int64_t i = rows[k];
        gather_family(s, &g, i, &fam[k], NULL);
        fcut[k] = (uint8_t*)malloc(fam[k].cnt ? fam[k].cnt : 1);
        for (int64_t e = 0; e < fam[k].cnt; ++e) {
            int64_t j = fam[k].v[e];
            int64_t a = i < j ? i : j, b = i < j ? j : i;
            fcut[k][e] = (uint8_t)orc_pair_cut(p, &s->X[3 * a], &s->X[3 * b]);
        }
    }
    for (int64_t k = 0; k < nrows; ++k) {
        need[rows[k]] = 1;
        for (int64_t e = 0; e < fam[k].cnt; ++e) need[fam[k].v[e]] = 1;
    }
    if (p->bdryeff_on) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) if (nthreads != 1)
#endif
        for (int64_t i = 0; i < s->n; ++i) {
            if (!need[i]) continue;
            ivec f = {0, 0, NULL};
            gather_family(s, &g, i, &f, NULL);
            s->be[i] = bdryeff_node(s, i, f.v, f.cnt, maxf);
            free(f.v);
        }
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) if (nthreads != 1)
#endif
    for (int64_t k = 0; k < nrows; ++k)
        row_force(s, rows[k], fam[k].v, fcut[k], fam[k].cnt, u, &F[3 * k], el ? &el[k] : NULL);
    grid_free(&g);
    for (int64_t k = 0; k < nrows; ++k) {
        free(fam[k].v);
        free(fcut[k]);
    }
    free(fam);
    free(fcut);
    free(need);
    orc_destroy(s);
    return 0;
}

/* CalcArealForceDensity + CalcForceOfPlusSpace (pdDataManager3.cpp:442-620), CCI branch, for
 * the current displacement U: the collinear nodes by exact coordinate equality, every node j
 * tested against the collinear node i with the rough 2 dx prefilter and CalcDistanceToCube
 * (i, j) < delta + tol, the half space (x_j - x_s) . e_n >= 0, j != i, the bond overload of
 * CalcBondForceCCI (pdDataManager2.cpp:1087-1154).  be (N, may be NULL) replaces bdryEff. */
void orc_areal_force_density(const orc_state* s, const int64_t* nodes, int64_t count, const double* be,
                             double* sigma)
{
    const orc_problem* p = &s->p;
    const double* bef = be ? be : s->be;
    const int ns = (int)(p->horizon / p->dx);
    const double ds = p->horizon / ns;
    const double tol = s->tol, dx = p->dx;
    for (int64_t t = 0; t < count; ++t) {
        const int64_t src = nodes[t];
        const double* xs = &s->X[3 * src];
        double sig[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int n = 0; n < 3; ++n) {
            /* collinear nodes in node-id order, as the reference collects them */
            for (int64_t i = 0; i < s->n; ++i) {
                const double* xi = &s->X[3 * i];
                int hit = 0;
                for (int is = 0; is != ns; ++is) {
                    if (n == 0 && xi[0] == xs[0] - dx * (1 + is) && xi[1] == xs[1] && xi[2] == xs[2]) hit = 1;
                    if (n == 1 && xi[0] == xs[0] && xi[1] == xs[1] - dx * (1 + is) && xi[2] == xs[2]) hit = 1;
                    if (n == 2 && xi[0] == xs[0] && xi[1] == xs[1] && xi[2] == xs[2] - dx * (1 + is)) hit = 1;
                }
                if (!hit) continue;
                double f[3] = {0.0, 0.0, 0.0};
                const orc_mat* mi = &p->mat[s->mat[i]];
                for (int64_t j = 0; j < s->n; ++j) {
                    const double* xj = &s->X[3 * j];
                    if (!(xj[n] - xs[n] >= 0.0) || j == i) continue;
                    double d = sqrt(sq(xj[0] - xi[0]) + sq(xj[1] - xi[1]) + sq(xj[2] - xi[2]));
                    if (d > p->horizon + 2.0 * dx) continue;
                    double dmin, dmax;
                    dist_to_cube(xi, xj, dx, &dmin, &dmax);
                    if (!(dmin < p->horizon + tol)) continue;
                    int a = i < j ? (int)i : (int)j, b = i < j ? (int)j : (int)i;
                    if (orc_pair_cut(p, &s->X[3 * a], &s->X[3 * b])) continue; /* slit extension */
                    const orc_mat* mj = &p->mat[s->mat[j]];
                    double xi3[3] = {xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]};
                    double eta[3] = {s->U[3 * j] - s->U[3 * i], s->U[3 * j + 1] - s->U[3 * i + 1],
                                     s->U[3 * j + 2] - s->U[3 * i + 2]};
                    double ecrit = (mj->ecrit < mi->ecrit) ? mj->ecrit : mi->ecrit;
                    double ci = s->spring[s->mat[i]], cj = s->spring[s->mat[j]];
                    double spring = (cj < ci) ? cj : ci;
                    double bij = 0.5 * (bef[i] + bef[j]);
                    double fb[3], el;
                    orc_bond_force_cci(p, xi3, eta, ecrit, spring, bij, 2.0, fb, &el);
                    f[0] += fb[0];
                    f[1] += fb[1];
                    f[2] += fb[2];
                }
                sig[3 * n] += f[0] * ds;
                sig[3 * n + 1] += f[1] * ds;
                sig[3 * n + 2] += f[2] * ds;
            }
        }
        memcpy(&sigma[9 * t], sig, sizeof sig);
    }
}

void orc_get_state(const orc_state* s, double* U, double* V)
{
    if (U) memcpy(U, s->U, sizeof(double) * 3 * s->n);
    if (V) memcpy(V, s->V, sizeof(double) * 3 * s->n);
}
/* U also re-interpolates the Gauss points of the non-BC nodes, as the end of a step does. */
void orc_set_state(orc_state* s, const double* U, const double* V)
{
    if (U) memcpy(s->U, U, sizeof(double) * 3 * s->n);
    if (U && s->ng) gauss_update(s, 0.0, s->U, 1);
    if (V) memcpy(s->V, V, sizeof(double) * 3 * s->n);
}
double orc_time(const orc_state* s) { return s->t; }
double orc_dt_current(const orc_state* s) { return s->dt_cur; }
double orc_ke(const orc_state* s) { return s->ke; }
double orc_ee(const orc_state* s) { return s->ee; }
int orc_steps_done(const orc_state* s) { return s->steps_done; }
int orc_num_gauss(const orc_state* s) { return s->ng; }
int64_t orc_num_contrib(const orc_state* s) { return s->cb_rp ? s->cb_rp[s->n] : 0; }
void orc_get_gauss(const orc_state* s, double* Xg, double* Ug)
{
    for (int64_t i = 0; i < s->n; ++i)
        for (int g = 0; g < s->ng; ++g)
            for (int c = 0; c < 3; ++c) {
                size_t q = ((size_t)i * s->ng + g) * 3 + c;
                if (Xg) Xg[q] = s->gofs[3 * g + c] + s->X[3 * i + c];
                if (Ug) Ug[q] = s->G[q];
            }
}
int64_t orc_num_ai_points(const orc_state* s) { return s->ai ? s->ai_ep[s->rp[s->n]] : 0; }
void orc_get_ai_points(const orc_state* s, int64_t* ent_ptr, double* xw)
{
    if (!s->ai) return;
    int64_t E = s->rp[s->n];
    if (ent_ptr) memcpy(ent_ptr, s->ai_ep, sizeof(int64_t) * (E + 1));
    if (xw) memcpy(xw, s->ai_xw, sizeof(double) * 4 * s->ai_ep[E]);
}
void orc_get_ai_mls(const orc_state* s, int64_t* po, double* phi)
{
    if (!s->ai_po) return;
    int64_t P = s->ai_ep[s->rp[s->n]];
    if (po) memcpy(po, s->ai_po, sizeof(int64_t) * (P + 1));
    if (phi) memcpy(phi, s->ai_phi, sizeof(double) * s->ai_po[P]);
}
void orc_get_mls(const orc_state* s, int64_t* cb_rp, int32_t* cb, double* phi)
{
    if (!s->cb_rp) return;
    int64_t m = s->cb_rp[s->n];
    if (cb_rp) memcpy(cb_rp, s->cb_rp, sizeof(int64_t) * (s->n + 1));
    if (cb) memcpy(cb, s->cb, sizeof(int32_t) * m);
    if (phi && s->phi) memcpy(phi, s->phi, sizeof(double) * (size_t)s->ng * m);
}
