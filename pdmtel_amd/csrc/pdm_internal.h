// pdm_internal.h -- handle layout, device records and kernel entry points.
//
// HBM layout (one specimen, N nodes, E directed family entries):
//   rec[2][N]  NodeRec, 64 B: reference position X (+ ecrit of the node's
//              material in .w) and the displacement every family gather reads
//              this stage (ping-pong: stage s reads rec[s&1], writes rec[(s+1)&1]).
//   row_ptr    int64 N+1, col int32 E (ascending within a row), cut uint8 E
//   coef64     FP64 E: bdryEff_ij * dx^3 * beta(r_ij) * c_ij, 0 for cut bonds
//   coef32     FP32 E: the same, rounded once (FP32-compensated path)
//   U,V,Vb,k1,k23  d4 per node (FP64 state; .w unused)
// The node order is the reference's lattice order (x fastest), so a family's
// neighbours sit in a few contiguous id runs and gathers hit L2.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "pdm.h"
#include "pdm_deck.h"

namespace pdm {

constexpr int kMaxMat = 8;
constexpr int kMaxBC = 16;
constexpr int kMaxSlit = 8;
constexpr int kMaxFamily = 1024;   // row-sort limit (families up to ~delta = 5.5 dx)

struct alignas(32) d4 {
    double x, y, z, w;
};

struct alignas(64) NodeRec {
    d4 x;   // position, .w = ecrit of the node's material
    d4 u;   // displacement seen by family gathers this stage
};

struct MatDev {
    double rho, g[3], ecrit, spring;
};

struct BCDev {
    int type;       // 1 DISPL, 2 VELOC, 3 DISPLGRAD, 4 TRACTION
    int f[3];
    int bcnode;     // 1: the node keeps U (reference BC-node treatment); 0: free node
    int pad;
    double v[9];
    double t_end, load;
    double trac[3]; // body-force density of a TRACTION bc (extension), 0 otherwise
};

struct SlitDev {
    int type;
    int axis;          // type 1: normal axis
    double p[11];      // type 1: box; type 2: centre(3) n(3) t(3) half_t half_s
    double w[3];       // type 2: n x t
};

struct StepParam {
    double t, dt;
};

// Arguments of the RK4 stage kernels.
struct StageArgs {
    const NodeRec* rin;
    NodeRec* rout;
    const int64_t* row_ptr;
    const int32_t* col;
    const double* coef64;
    const float* coef32;
    const int32_t* mat;
    const int32_t* bc;
    const int32_t* nb;      // unbroken family entries per node
    d4* U;
    d4* V;
    d4* Vb;
    d4* k1;
    d4* k23;
    double* elast;          // per-node elastic energy density, summed over the 4 stages
    double* part;           // per-block partial KE / EE (stage 4)
    const StepParam* steps; // per-step t, dt
    int64_t n;
    double damping;
    double vol;             // pow(dx, 3): node volume for energies
    double inv_vol_cube;    // 1 / (dx*dx*dx): coef -> bdryEff * beta * c for energies
    int nmat, nbc;
    int energies;
    int xcd_remap;
    MatDev mats[kMaxMat];
    BCDev bcs[kMaxBC];
};

// Arguments of the standalone force evaluation (pdm_compute_forces / tangent).
struct ForceArgs {
    const NodeRec* rin;
    const int64_t* row_ptr;
    const int32_t* col;
    const double* coef64;
    const float* coef32;
    const int32_t* mat;
    double* F;      // 3N AoS output
    double* elast;  // N or null
    int64_t n;
    double inv_vol_cube;
    MatDev mats[kMaxMat];
};

// Bond-build parameters.
struct BondArgs {
    const d4* X;
    int64_t n;
    double dx, horizon, tol;
    double lo[3];
    double h;              // cell edge
    int64_t nc[3];
    int nslit;
    SlitDev slits[kMaxSlit];
};

}  // namespace pdm

struct pdm_handle {
    pdm_options opt{};
    pdm::Deck deck;
    std::string err;
    bool inited = false;
    int dev = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    // host copies
    int64_t n = 0, E = 0;
    std::vector<double> X;          // 3N AoS
    std::vector<int32_t> mat, bc;
    std::vector<double> spring, sound;
    double min_bond = 0, max_c = 0, dt_stable = 0;
    int max_family = 0;
    double trac_density[pdm::kMaxBC][3] = {};

    // device
    pdm::NodeRec* rec[2] = {nullptr, nullptr};
    pdm::d4* Xd = nullptr;
    int32_t *matd = nullptr, *bcd = nullptr, *nb = nullptr, *col = nullptr;
    int64_t* rp = nullptr;
    uint8_t* cut = nullptr;
    double* coef64 = nullptr;
    float* coef32 = nullptr;
    double* be = nullptr;
    pdm::d4 *U = nullptr, *V = nullptr, *Vb = nullptr, *k1 = nullptr, *k23 = nullptr;
    double* elast = nullptr;
    double* part = nullptr;        // per-block partial energies
    double* energy = nullptr;      // per-step KE, EE (chunk)
    pdm::StepParam* steps = nullptr;
    double* scratchF = nullptr;    // 3N FP64 for standalone forces
    int64_t part_blocks = 0;
    int chunk_cap = 0;

    // dynamic clock (mirrors pdSolve::SolveDynamic's loop variables)
    double t = 0, dt_cur = 0;
    int step = 0, last_step = 0;
    double ke = 0, ee = 0;
    double last_kernel_ms = 0;
    int64_t last_kernel_launches = 0;

    // graph cache
    hipGraphExec_t graph_exec = nullptr;
    int graph_steps = 0;
    int graph_rec_parity = -1;
};

namespace pdm {

// host model setup (pdm_model.cpp)
int build_nodes(pdm_handle* h);
void init_materials(pdm_handle* h);
int count_max_family(double horizon, double dx, double tol);
double bdryeff_type_one(const Deck& d, double spring, double E, double h);
double time_factor(const BoundaryCondition& b, double t);
bool is_bc_node(const Deck& d, int bcid);
void fill_stage_tables(const pdm_handle* h, StageArgs& a);

// device setup (pdm_bonds.hip)
int build_bonds(pdm_handle* h);
int build_bdryeff_and_coef(pdm_handle* h);

// kernels (pdm_kernels.hip)
hipError_t launch_stage(int stage, int precision, const StageArgs& a, int step_idx, hipStream_t s);
hipError_t launch_energy_final(const StageArgs& a, int64_t nblocks, double* out2, hipStream_t s);
hipError_t launch_forces(int precision, const ForceArgs& a, hipStream_t s);
hipError_t launch_tangent(int precision, const ForceArgs& a, const double* u3, double* y3, hipStream_t s);
int64_t stage_blocks(int64_t n);

}  // namespace pdm
