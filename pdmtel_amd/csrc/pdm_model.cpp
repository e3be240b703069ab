// pdm_model.cpp -- host-side model setup: nodes, materials, time factors.
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>

#include "pdm_geom.h"
#include "pdm_internal.h"

namespace pdm {
namespace {

bool in_space(const Space& s, const double* x, double tol)
{
    const double* b = s.b;
    if (s.geo == 1)
        return x[0] >= b[0] - tol && x[0] <= b[1] + tol && x[1] >= b[2] - tol && x[1] <= b[3] + tol &&
               x[2] >= b[4] - tol && x[2] <= b[5] + tol;
    if (s.geo == 2) {
        double dist = std::sqrt((x[0] - b[1]) * (x[0] - b[1]) + (x[1] - b[2]) * (x[1] - b[2]));
        return dist <= b[0] + tol && x[2] >= b[3] - tol && x[2] <= b[4] + tol;
    }
    return false;
}

// external grid deletion has no tolerance (pdDataManager3.cpp:836-867)
bool in_space_exact(const Space& s, const double* x) { return in_space(s, x, 0.0); }

uint64_t splitmix64(uint64_t& st)
{
    uint64_t z = (st += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

}  // namespace

// InitNode (pdDataManager.cpp:813-1098).  Lattice nodes in id order z -> y -> x (x fastest),
// deletion spaces with tolerance 1e-4 dx, then the material of the LAST material space
// (a node outside a later space loses its material, pdDataManager.cpp:918-968) and the
// boundary condition of the last boundary space containing the node.
// grid_file: the reference reads nodes but never stores them (pdDataManager3.cpp:827-877);
// here they are kept, with the file's node order as ids.
// grid_jitter (extension): SplitMix64(seed), node-id order, x then y then z, U(-a, a) dx.
int build_nodes(pdm_handle* h)
{
    const Deck& d = h->deck;
    double tol = d.tol;
    h->X.clear();
    if (!d.grid_file.empty()) {
        std::ifstream in(d.grid_file.c_str());
        if (!in) {
            h->err = "grid_file " + d.grid_file + " cannot be opened";
            return PDM_E_IO;
        }
        long long cnt = 0;
        in >> cnt;
        for (long long i = 0; i < cnt; ++i) {
            double x[3];
            int mid = 0;
            if (!(in >> x[0] >> x[1] >> x[2] >> mid)) {
                h->err = "grid_file " + d.grid_file + " ends early";
                return PDM_E_PARSE;
            }
            bool del = false;
            for (const auto& s : d.del_spaces)
                if (in_space_exact(s, x)) del = true;
            if (!del) h->X.insert(h->X.end(), x, x + 3);
        }
    } else {
        for (int n = 1; n != d.nz + 1; ++n)
            for (int m = 1; m != d.ny + 1; ++m)
                for (int l = 1; l != d.nx + 1; ++l) {
                    double x[3];
                    x[0] = (l - (d.nx + 1) * 0.5) * d.dx + d.cx;
                    x[1] = (m - (d.ny + 1) * 0.5) * d.dx + d.cy;
                    x[2] = (n - (d.nz + 1) * 0.5) * d.dx + d.cz;
                    bool del = false;
                    for (const auto& s : d.del_spaces)
                        if (in_space(s, x, tol)) del = true;
                    if (!del) h->X.insert(h->X.end(), x, x + 3);
                }
    }
    h->n = (int64_t)(h->X.size() / 3);
    if (h->n >= (int64_t)INT32_MAX) {
        h->err = "node count exceeds the int32 column index range";
        return PDM_E_LIMIT;
    }
    if (d.jitter != 0.0) {
        uint64_t st = d.seed;
        for (int64_t i = 0; i < h->n; ++i)
            for (int c = 0; c < 3; ++c) {
                double u01 = (double)(splitmix64(st) >> 11) * 0x1.0p-53;
                h->X[3 * i + c] += (2.0 * u01 - 1.0) * d.jitter * d.dx;
            }
    }
    h->mat.assign(h->n, -1);
    h->bc.assign(h->n, -1);
    for (int64_t i = 0; i < h->n; ++i) {
        const double* x = &h->X[3 * i];
        int m = -1;
        for (const auto& s : d.mat_spaces) m = in_space(s, x, tol) ? s.ref : -1;
        if (m < 0) {
            std::ostringstream os;
            os << "InitNode: node " << i << " not in any material space! (x1 = " << x[0] << ", x2 = " << x[1]
               << ", x3 = " << x[2] << ")";
            h->err = os.str();
            return PDM_E_MODEL;
        }
        h->mat[i] = m;
        int b = -1;
        for (const auto& s : d.bc_spaces)
            if (in_space(s, x, tol)) b = s.ref;
        h->bc[i] = b;
    }
    return PDM_OK;
}

// InitMaterial (pdDataManager.cpp:1297-1338): c = 18 K / (pi delta^4), K = E / (3 (1 - 2 nu)).
void init_materials(pdm_handle* h)
{
    const Deck& d = h->deck;
    const double pi = std::acos(-1.0);
    h->spring.assign(d.mats.size(), 0.0);
    h->sound.assign(d.mats.size(), 0.0);
    h->max_c = 0.0;
    for (size_t m = 0; m < d.mats.size(); ++m) {
        const Material& mt = d.mats[m];
        double K = mt.E / (3.0 * (1.0 - 2.0 * mt.nu));
        h->spring[m] = (mt.type == 1) ? 18 * K / (pi * std::pow(d.horizon, 4.0)) : 0.0;
        h->sound[m] = std::sqrt(K / mt.rho);
        h->max_c = (h->max_c < h->sound[m]) ? h->sound[m] : h->max_c;
    }
}

// Family count of a full horizon on a virtual lattice (pdDataManager.cpp:1345-1387).
int count_max_family(double horizon, double dx, double tol)
{
    int vd = 2 * ((int)(horizon / dx) + 1);
    int cnt = 0;
    for (int n = 1; n != vd + 1; ++n)
        for (int m = 1; m != vd + 1; ++m)
            for (int l = 1; l != vd + 1; ++l) {
                double x1 = (l - vd * 0.5) * dx, x2 = (m - vd * 0.5) * dx, x3 = (n - vd * 0.5) * dx;
                double dmin = g_dist_to_cube(0.0, 0.0, 0.0, x1, x2, x3, dx);
                if (dmin < horizon + tol) ++cnt;
            }
    return cnt;
}

// CalcBdryEffFactorTypeOne (pdDataManager2.cpp:1479-1485)
double bdryeff_type_one(const Deck& d, double spring, double E, double h)
{
    const double pi = std::acos(-1.0);
    double full = 3.0 * E;
    double part = full - 0.5 * spring * pi *
                             (std::pow(d.horizon, 4) / 4.0 - h * std::pow(d.horizon, 3) / 3.0 + std::pow(h, 4) / 12.0);
    return full / part;
}

// GetTimeFactor (pdBdryConditionDispl.cpp:15-27; DisplGrad and Tract identical; Veloc :15-18)
double time_factor(const BoundaryCondition& b, double t)
{
    if (b.type == 2) return (t <= b.t_end) ? 1.0 : 0.0;
    if (b.load == 1.0) return (t <= b.t_end) ? t / b.t_end : 1.0;
    return 1.0;
}

// A node in a boundary space keeps its U through the stages (pdDataManager.cpp:1831-1854).
// With traction_as_body_force, TRACTION nodes are free nodes instead: the reference's
// stale-U treatment of a type-4 node has amplification determinant 1 + 0.875 (w dt)^2.
bool is_bc_node(const Deck& d, int bcid)
{
    if (bcid < 0) return false;
    if (d.traction_body && d.bcs[bcid].type == 4) return false;
    return true;
}

void fill_spec_tables(const pdm_handle* h, SpecDev& sp, MatDev* mats, BCDev* bcs)
{
    const Deck& d = h->deck;
    std::memset(mats, 0, sizeof(MatDev) * kMaxMat);
    std::memset(bcs, 0, sizeof(BCDev) * kMaxBC);
    const int nmat = (int)d.mats.size(), nbc = (int)d.bcs.size();
    for (int m = 0; m < nmat; ++m) {
        mats[m].rho = d.mats[m].rho;
        for (int c = 0; c < 3; ++c) mats[m].g[c] = d.mats[m].g[c];
        mats[m].ecrit = d.mats[m].ecrit;
        mats[m].spring = h->spring[m];
    }
    for (int b = 0; b < nbc; ++b) {
        BCDev& o = bcs[b];
        const BoundaryCondition& s = d.bcs[b];
        std::memset(&o, 0, sizeof(o));
        o.type = s.type;
        for (int c = 0; c < 3; ++c) o.f[c] = s.f[c];
        o.bcnode = is_bc_node(d, b) ? 1 : 0;
        for (int k = 0; k < 9; ++k) o.v[k] = s.v[k];
        o.t_end = s.t_end;
        o.load = s.load;
        for (int c = 0; c < 3; ++c) o.trac[c] = h->trac_density[b][c];
    }
    sp.ecrit0 = d.mats.empty() ? 0.0 : d.mats[0].ecrit;
    sp.inv_vol_cube = 1.0 / (d.dx * d.dx * d.dx);
    sp.fscale = h->fscale;
    sp.ivc4 = 0.25 * sp.inv_vol_cube;
    sp.ivc4_rec = 0.25 * sp.inv_vol_cube * (1.0 / sp.fscale);
    sp.damping = d.damping;
    sp.vol = std::pow(d.dx, 3.0);
    sp.kc0 = h->spring.empty() ? 0.0 : 0.5 * (d.dx * d.dx * d.dx) * h->spring[0];
    // beta (pdDataManager3.cpp:1801-1831) in the FP32 record units (m x fscale)
    // (entries with beta 0, including original CCI's beyond delta, read the padding record)
    sp.beta0 = (float)((d.horizon + 0.5 * d.dx) / d.dx - 128.0 * 0x1p-26);
    sp.beta1 = (float)(1.0 / (d.dx * h->fscale));
}

}  // namespace pdm
