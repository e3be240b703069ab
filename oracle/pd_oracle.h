/*
 * pd_oracle.h -- CPU oracle for the linear bond-based peridynamics hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64 restatement of the
 * reference solver's algorithms, used as the checker for the HIP library.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library never links, calls or falls back to it.
 *
 * Every function states the reference file:line it restates.  Arrays that
 * hold 3-vectors are node-major AoS (x,y,z of node 0, then node 1, ...),
 * i.e. the reference's displStep/_rowOfNode layout (pdDataManager.cpp:1085).
 *
 * Pinning (what checks this restatement independently of its own formulas)
 * is listed per function in the header of pd_oracle.c.
 */
#ifndef PD_ORACLE_H
#define PD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_MAT 8
#define ORC_MAX_BC 16
#define ORC_MAX_SPACE 16
#define ORC_MAX_SLIT 8

typedef struct {
    int type;          /* 1 bond-based (PMB), 2 state-based (never used by RK4) */
    double E, sigy, ecrit, rho, nu;
    double g[3];       /* gravity_N */
} orc_mat;

typedef struct {
    int type;          /* 1 DISPL, 2 VELOC, 3 DISPLGRAD, 4 TRACTION */
    int f[3];          /* direction flags */
    double v[9];       /* 3 values (types 1,2,4) or 9 gradient values (type 3) */
    double t_end, load;/* end time, load type (1.0 ramp, else step) */
} orc_bc;

typedef struct {
    int geo;           /* 1 cuboid: b = x1lo x1hi x2lo x2hi x3lo x3hi; 2 cylinder: b = r cx cy x3lo x3hi */
    int ref;           /* material index / bc index (0-based); unused for deletion */
    double b[6];
} orc_space;

typedef struct {
    int type;          /* 1 axis-normal rectangle (lo/hi box with one zero extent), 2 oriented rectangle */
    double p[11];      /* type 1: x1lo x1hi x2lo x2hi x3lo x3hi; type 2: c(3) n(3) t(3) half_t half_s */
} orc_slit;

typedef struct {
    int nx, ny, nz;
    double dx, horizon, cx, cy, cz;
    double total_time;
    int total_step, dump;
    double safety;
    int fixed_dt_on;
    double fixed_dt;
    int nmat; orc_mat mat[ORC_MAX_MAT];
    int nbc; orc_bc bc[ORC_MAX_BC];
    int nms; orc_space ms[ORC_MAX_SPACE];
    int nds; orc_space ds[ORC_MAX_SPACE];
    int nbs; orc_space bs[ORC_MAX_SPACE];
    double damping;
    int bdryeff_on;
    int cci_on, cci_mdf, fgi_on, intg_point, mls_on;
    double rw;
    /* extensions the reference lacks (declared deviations) */
    double jitter;            /* uniform jitter amplitude as a fraction of dx */
    uint64_t seed;            /* SplitMix64 seed for the jitter */
    int nslit; orc_slit slit[ORC_MAX_SLIT];
    int traction_body;        /* 1: TRACTION BCs act as a body force (total force / face nodes) */
    int ai_on;                /* adaptive integration (cci_on 0, fgi_on 0, ai_on 1 mls_on): intg_point is
                                 the trapezoidal index, 2^(n-1)+1 points per direction */
} orc_problem;

typedef struct orc_state orc_state;

orc_state* orc_create(const orc_problem* p);
void orc_destroy(orc_state* s);
/* InitNode + InitBond + InitMaterial + InitBdryEffectFactor.
 * bond_method 0: the reference O(N^2) pair loop; 1: cell list, same predicate. */
int orc_init(orc_state* s, int bond_method, int nthreads);
int64_t orc_num_nodes(const orc_state* s);
int64_t orc_num_entries(const orc_state* s);      /* directed family entries = 2 x bonds */
int orc_num_max_family(const orc_state* s);
double orc_min_bond_length(const orc_state* s);
double orc_stable_dt(const orc_state* s);
double orc_spring(const orc_state* s, int mat);
void orc_get_nodes(const orc_state* s, double* X, int* mat, int* bc);
void orc_get_csr(const orc_state* s, int64_t* row_ptr, int32_t* col, uint8_t* cut);
void orc_get_bdryeff(const orc_state* s, double* be);
void orc_set_bdryeff(orc_state* s, const double* be);
/* Family force sum F_i = sum_j f_ij (no body force) for the displacement u (3N AoS),
 * CCI path of CalcBondBasedMotionRK4 with every node's u taken from u. elast (N) may be NULL. */
void orc_forces(const orc_state* s, const double* u, double* F, double* elast, int nthreads);
/* Linearised (tangent at eta=0) bond operator y = K u, same weights as the force. */
void orc_apply_tangent(const orc_state* s, const double* u, double* y, int nthreads);
/* Static system: free mask, prescribed values, loads (3N each). */
void orc_static_system(const orc_state* s, double* mask, double* uc, double* b);
/* nsteps of pdSolve::SolveDynamic's loop (t advanced before RK4, shortened last step). */
int orc_run_dynamic(orc_state* s, int nsteps, int nthreads);
void orc_get_state(const orc_state* s, double* U, double* V);
void orc_set_state(orc_state* s, const double* U, const double* V);
double orc_time(const orc_state* s);
double orc_dt_current(const orc_state* s);
double orc_ke(const orc_state* s);
double orc_ee(const orc_state* s);
int orc_steps_done(const orc_state* s);

/* Fixed Gaussian integration (fgi_on, cci_on 0): n^3 Gauss points per node.
 * orc_gauss_rule: the order-n Gauss-Legendre rule in the reference's point order (-1: no rule).
 * orc_get_gauss: point positions and displacements, [node][point][xyz].
 * orc_get_mls: contributing nodes (row pointer N+1, ids) and coefficients, node i's block
 * [point][contributing node] at num_gauss * cb_rp[i]. */
int orc_gauss_rule(int n, double* a, double* w);
int orc_num_gauss(const orc_state* s);
int64_t orc_num_contrib(const orc_state* s);
void orc_get_gauss(const orc_state* s, double* Xg, double* Ug);
void orc_get_mls(const orc_state* s, int64_t* cb_rp, int32_t* cb, double* phi);

/* Adaptive integration (ai_on): trapezoidal points of each family entry over node j's cell
 * clipped to node i's horizon, in family coordinates (node i at the origin).
 * orc_ai_cell_points: the points (x, y, z, weight; 4 per point) of one cell whose centre is c
 * (X_j - X_i) for the bond distances dmin / dmax; returns the count (points beyond cap are
 * counted, not written).
 * orc_get_ai_points: ent_ptr (E + 1, reference CSR entry order) and the points (4 P).
 * orc_get_ai_mls: the MLS coefficients of each point over node j's contributing nodes
 * (orc_get_mls's cb_rp / cb), point q's at phi[po[q] .. po[q + 1]). */
int orc_ai_cell_points(const orc_problem* p, const double* c, double dmin, double dmax, double* xw, int cap);
/* The integrated bond force of one such cell for the relative displacement eta = G xi (G 3x3
 * row-major), with bdryEff_ij, spring constant and ecrit given; returns the point count. */
int orc_ai_cell_force(const orc_problem* p, const double* c, double dmin, double dmax, const double* G, double bij,
                      double spring, double ecrit, double* F, double* el);
int64_t orc_num_ai_points(const orc_state* s);
void orc_get_ai_points(const orc_state* s, int64_t* ent_ptr, double* xw);
void orc_get_ai_mls(const orc_state* s, int64_t* po, double* phi);

/* Families (ascending), cut flags and bdryEff factors of the given rows only (for problems
 * too large to build whole): rp (nrows + 1), col / cut (cap entries), be (nrows), rows in the
 * given order.  Returns the entry count, -1 if it exceeds cap, or a negative init code. */
int64_t orc_sample_rows(const orc_problem* p, const int64_t* rows, int64_t nrows, int64_t* rp, int32_t* col,
                        uint8_t* cut, double* be, int64_t cap, int nthreads);

/* Family force sums (CCI, no body force) and energy densities of the given rows only, for a
 * displacement field u of every node (3N): F (3 nrows), el (nrows, may be NULL). */
int orc_sample_forces(const orc_problem* p, const int64_t* rows, int64_t nrows, const double* u, double* F,
                      double* el, int nthreads);

/* Areal force density (CalcArealForceDensity, CCI) at the given nodes from the current U:
 * sigma[9 t + 3 n + c]; be (N, may be NULL) replaces the boundary-effect factors. */
void orc_areal_force_density(const orc_state* s, const int64_t* nodes, int64_t count, const double* be,
                             double* sigma);

/* 0: reference pow(x, 2.0) via glibc (default); 1: correctly rounded square x*x */
void orc_set_square_mode(int exact);

/* single-pair primitives exposed for tests */
int orc_pair_bonded(const orc_problem* p, const double* xi, const double* xj, double* dmin, double* dmax);
int orc_pair_cut(const orc_problem* p, const double* xi, const double* xj);
double orc_beta(double dist, double delta, double dx, int cci_on, int cci_mdf);
void orc_bond_force_cci(const orc_problem* p, const double* xi, const double* eta, double ecrit,
                        double spring, double bdryEffij, double effWeight, double* f, double* elast);
double orc_time_factor(const orc_bc* bc, double t);

#ifdef __cplusplus
}
#endif
#endif
