// pdm_internal.h -- handle layout, device records and kernel entry points.
//
// HBM layout of one specimen (N nodes, E directed family entries, T tiles):
//
//  Reference order (built once, kept for pdm_get_csr / bdryEff output):
//    Xd d4[N], rp int64[N+1], col int32[E] (ascending per row), cut uint8[E], be FP64[N]
//
//  Internal (tile-major) order, read every RK4 stage:
//    Nodes are binned into spatial tiles (edge >= the family reach along each
//    axis is not required: the neighbourhood spans R tiles per axis) and renumbered
//    tile by tile.  For tile t the union of its nodes' families (and the nodes
//    themselves) is the halo list halo[halo_ptr[t] .. halo_ptr[t+1]), sorted by
//    internal id; each family entry stores its neighbour's position in that list
//    as a uint16 slot.  A workgroup stages the halo's X and u (FP64) in LDS once,
//    then every family gather is an LDS read.
//    Xi d4[N]      position, .w = ecrit of the node's material
//    u[2] d4[N]    displacement gathered by the next stage (ping-pong)
//    rpi int64[N+1], slot uint16[Ei], coef64 FP64[Ei] / coef32 FP32[Ei]
//      CCI rows hold the entries with beta > 0 and no cut (Elive of them; the others
//      carry no force), fixed-Gaussian rows every entry; rows padded to a multiple of
//      kRowAlign entries; padding entries and fixed-Gaussian cut bonds point at the
//      tile's padding record, with coef 0
//      coef = bdryEff_ij * dx^3 * beta(r_ij) * c_ij
//    bk BK[N]      bdryEff and dx^3 c / 2 in FP32, and corr uint8[Ei]: the FP32 force
//                  kernels form each entry's coefficient from these and beta(r), corrected
//                  per entry by (corr - 128) 2^-26, instead of reading coef32
//    U, V, Vb, k1, k23 d4[N]  FP64 RK4 state; mat, bc, nb int32[N]
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
constexpr int kMaxFamily = 1024;     // row-sort limit (families up to ~delta = 5.5 dx)
constexpr int kMaxTileNodes = 343;   // nodes per tile (tile edge up to 7 dx)
constexpr size_t kLdsBytes = 160 * 1024;  // LDS per CU on gfx950
constexpr int kMaxFlags = 8192;      // nodes in a tile's neighbourhood (halo build)
constexpr int kMaxHalo = 2560;       // halo entries staged in LDS: 7 x 8 B x 2560 + 32 B x 512 < 160 KB
constexpr int kRowAlign = 8;          // internal rows padded to whole blocks of 8 entries
constexpr int64_t kNoAllocMaxNodes = 1 << 19;  // nodes pdm_compute_forces_noalloc's scratch holds

struct alignas(32) d4 {
    double x, y, z, w;
};

// Per node, FP32: the boundary-effect factor and half the node volume times the spring
// constant, from which the FP32 force kernels form coef = (be_i + be_j) min(kc_i, kc_j) beta.
struct alignas(8) BK {
    float be, kc;
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

// Constants of one specimen.  A launch covers the tiles of one or many specimens;
// each tile reads its specimen's entry (tile_spec), and its material and BC tables at
// mats[spec * kMaxMat], bcs[spec * kMaxBC].
struct SpecDev {
    double ecrit0;        // common ecrit of the specimen's materials (when uniform_mat)
    double inv_vol_cube;  // 1 / dx^3
    double fscale;        // fixed-point scale of the FP32 halo records
    double damping;       // damping_coeff
    double vol;           // pow(dx, 3): node volume for the energies
    double kc0;           // dx^3 c / 2 of the specimen's materials (when uniform_mat)
    double ivc4;          // 0.25 / dx^3: energy-density scale of the FP64 family sums
    double ivc4_rec;      // 0.25 / dx^3 / fscale: energy-density scale of the FP32 record sums
    // beta(r) of the FP32 records (r in record units, S m), offset by the correction bias:
    // med3(beta0 - r beta1, -128 2^-26, 1 - 128 2^-26)
    float beta0, beta1;
};

// Tile-gather inputs shared by every tiled kernel.
struct TileView {
    const int32_t* tiles;      // launch index -> tile id
    const int64_t* tile_start; // tile id -> first internal node (size ntiles + 1)
    const int64_t* halo_ptr;   // launch index -> halo list start (size ntl + 1)
    const int32_t* halo;       // internal ids
    const uint16_t* self_slot; // internal node -> its slot in its own tile's halo
    const int64_t* rpi;
    const uint16_t* slot;
    const double* coef64;
    const float* coef32;
    const BK* bk;
    const uint8_t* corr;
    const d4* X;
    const d4* tref;            // launch index -> tile centre (reference point of FP32 records)
    int64_t n;
    int64_t ntl;               // launched (non-empty) tiles
    int halo_max;              // largest halo (sizes the LDS staging area)
    int tile_max;              // most nodes in a tile (sizes the LDS result area)
    int xcd_remap;
    int waves;                 // wavefronts per tile workgroup (4 or 8)
    int fam_per_wave;          // families a wavefront sums at once (1, 2 or 4)
    int uniform_mat;           // each specimen's materials share one ecrit and one spring constant:
                               // no per-neighbour ecrit in LDS, and kc0 scales the FP32 family sums
    const SpecDev* specs;      // per-specimen constants
    const int32_t* tile_spec;  // launch index -> specimen (nullptr: one specimen)
    const int16_t* wlo;        // PDM_FP32C2 tangent weights: 16-bit correction of each coef32
    const int32_t* lmap;       // stage launch over a class of tiles: workgroup -> launch index
    int64_t nmap;              // workgroups of such a launch (lmap nullptr: every tile, ntl)
};

// A batch's tiles whose LDS footprint allows the same number of tiles per CU, launched
// together (one stage launch per class): specimens with small halos then run at their own
// occupancy instead of that of the batch's largest halo.
struct TileClass {
    int32_t* map = nullptr;    // class workgroup -> launch index (device)
    int64_t n = 0;             // tiles in the class
    int halo_max = 0, tile_max = 0;
    int waves = 8, fams = 2;   // layout of the class's launches
    double tune_ms = 0.0;      // stage time of that layout when autotuned
};

// Arguments of the RK4 stage kernels.
struct StageArgs {
    TileView tv;
    const d4* uin;
    d4* uout;
    const int32_t* mat;
    const int32_t* bc;
    const int32_t* nb;      // unbroken family entries per node
    d4* U;
    d4* V;
    d4* Vb;
    d4* k1;
    d4* k23;
    double* elast;          // per-node elastic energy density, summed over the 4 stages
    double* part;           // per-step, per-tile partial KE / EE of stage 4: [step][tile][2]
    const StepParam* steps; // [step][specimen] t, dt
    const MatDev* mats;     // [specimen][kMaxMat]
    const BCDev* bcs;       // [specimen][kMaxBC]
    int nspec;
    int energies;
};

// Fixed Gaussian integration (fgi_on): ng = n^3 Gauss points per node, internal order.
// tab[8 g ..]: the point's offset (0.5 dx a) from its node in x, y, z; its force weight
// dx^3/8 w1 w2 w3; its beta cell size (w1 w2 w3)^(1/3) dx / 2; its energy factor
// ((w1 w2 w3)^(1/3) / 2)^3.
constexpr int kFgiTab = 8;

struct FgiArgs {
    const double* tab;      // [ng][kFgiTab]
    int ng;
    int mode;               // k_fgi_gauss: 0 as the reference, 1 non-BC nodes only, 2 every node
    double horizon;
    const d4* gin;          // Gauss-point displacements [N][ng] read by the family sums
    d4* gout;               // written by k_fgi_gauss
    const int64_t* cb_ptr;  // MLS contributing nodes (n > 1): [N + 1]
    const int32_t* cb;
    const double* phi;      // coefficient k of point g of node i at (cb_ptr[i] + k) * ng + g
    const d4* X;
    const int32_t* bc;
    const BCDev* bcs;
    const StepParam* steps; // the step's time (mode 0)
    int64_t n;
    d4* trial;              // RK4 trial displacement of every node: two buffers of n, stage s writes
                            // the one at (s & 1) n, which stage s + 1 reads
    d4* u0;                 // U at the start of the step (BC nodes)
    const int64_t* pair_ptr;  // [N + 1] live (entry, Gauss point) pairs of each internal row
    const uint2* pairs;       // ((entry offset in the row) << 16 | point, node j's internal id)
    int prec;                 // fixed Gaussian bond arithmetic: PDM_FP64, PDM_FP32C, PDM_FP32C2
    // adaptive integration (ai_on): trapezoidal points of each internal row's entries
    int ai;                   // 0 off, 1 MLS point displacements, 2 relative displacement G_j xi
    const int64_t* ai_ptr;    // [N + 1]
    const d4* ai_xw;          // [P] position relative to node i (x, y, z) and weight
    const uint32_t* ai_eo;    // [P] entry offset in node i's internal row
    const int64_t* ai_po;     // [P] the point's first MLS coefficient in ai_phi, coefficient k at
    const int32_t* ai_ps;     //     ai_po + k * ai_ps (an entry's block is [contributing node][point])
    const double* ai_phi;     // coefficients over node j's contributing nodes (cb_ptr / cb)
    const int32_t* ai_i;      // [P] the point's nodes i and j (internal ids)
    const int32_t* ai_j;
    const d4* ai_res;         // [N] family sums of k_ai_family, read by the stage kernel
    int64_t ai_np;
    double ai_vol;            // dx^3: a point's energy share is its weight / dx^3
};

// Device temporaries of a set-up routine, released on every return path once the
// stream has drained: register each pointer (still null) before allocating it.
struct DevTemps {
    hipStream_t s;
    std::vector<void**> p;
    explicit DevTemps(hipStream_t st) : s(st) {}
    template <class T>
    void add(T** q) { p.push_back(reinterpret_cast<void**>(q)); }
    ~DevTemps()
    {
        (void)hipStreamSynchronize(s);
        for (void** q : p)
            if (*q) {
                (void)hipFree(*q);
                *q = nullptr;
            }
    }
    DevTemps(const DevTemps&) = delete;
    DevTemps& operator=(const DevTemps&) = delete;
};

// Areal force density (pdm_post.hip): reference-order node data and the bond-force constants.
struct AfdArgs {
    const d4* X;            // positions
    const d4* U;            // displacements
    const int64_t* rp;      // reference-order CSR
    const int32_t* col;
    const uint8_t* cut;
    const int32_t* mat;
    const double* be;       // boundary-effect factors
    double ecrit[kMaxMat], spring[kMaxMat];
    double horizon, dx, vol, ds;
    int ns;                 // collinear nodes per direction, int(delta / dx)
    int original_cci;
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

    // host copies (reference order)
    int64_t n = 0, E = 0;
    int64_t Ei = 0;                 // internal (padded) family entries
    int64_t Elive = 0;              // entries the gathers evaluate (CCI: uncut, beta > 0)
    std::vector<double> X;          // 3N AoS
    std::vector<int32_t> mat, bc;
    std::vector<int32_t> perm;      // internal -> reference id
    std::vector<double> spring, sound;
    double min_bond = 0, max_c = 0, dt_stable = 0;
    int max_family = 0;
    double trac_density[pdm::kMaxBC][3] = {};

    // device, reference order
    pdm::d4* Xd = nullptr;
    int64_t* rp = nullptr;
    int32_t* col = nullptr;
    uint8_t* cut = nullptr;
    double* be = nullptr;
    int32_t* matd = nullptr;

    // device, internal (tile-major) order
    int32_t* permd = nullptr;
    pdm::d4* Xi = nullptr;
    pdm::d4* u[2] = {nullptr, nullptr};
    int64_t* rpi = nullptr;
    uint16_t* slot = nullptr;
    uint16_t* self_slot = nullptr;
    double* coef64 = nullptr;
    float* coef32 = nullptr;
    pdm::BK* bk = nullptr;
    uint8_t* corr = nullptr;
    int32_t *mati = nullptr, *bci = nullptr, *nb = nullptr;
    int32_t* tiles = nullptr;
    int64_t* tile_start = nullptr;
    pdm::d4* tref = nullptr;            // launch index -> tile centre
    pdm::SpecDev* specs_d = nullptr;    // per-specimen constants (one entry for a single specimen)
    pdm::MatDev* mats_d = nullptr;      // [specimen][kMaxMat]
    pdm::BCDev* bcs_d = nullptr;        // [specimen][kMaxBC]
    int32_t* tile_spec_d = nullptr;     // launch index -> specimen (batches only)
    std::vector<pdm::TileClass> tclass; // batches: stage launches per class (empty: one launch)
    std::vector<int> spec_class;        // batches with classes: specimen -> class
    int64_t* spec_tile_ptr_d = nullptr; // specimen -> first launch index (nspec + 1)
    int nspec = 1;
    std::vector<pdm_handle*> members; // batch: the specimens (parsed decks, nodes, bonds, clocks)
    std::vector<int64_t> spec_node_ptr; // batch: specimen -> first node of the combined arrays
    bool shared_stream = false;     // a batch member uses its batch's stream
    std::vector<double> last_energy; // [spec][2] KE, EE of each specimen's last step
    double fscale = 1.0;                // fixed-point scale of the FP32 halo records
    int64_t* halo_ptr = nullptr;
    int32_t* halo = nullptr;
    int64_t ntiles = 0, ntl = 0, halo_total = 0;
    int halo_max = 0, tile_max = 0;
    int tile_edge[3] = {4, 4, 4};   // tile edge in grid spacings
    int waves = 8;                  // wavefronts per tile workgroup
    int fam_per_wave = 2;           // families a wavefront sums at once
    int tuned = 0;                  // the layout was chosen by timing candidates at pdm_init
    double tune_ms = 0.0;           // stage launch time of the chosen layout (ms)
    int uniform_mat = 1;

    pdm::d4 *U = nullptr, *V = nullptr, *Vb = nullptr, *k1 = nullptr, *k23 = nullptr;
    double* elast = nullptr;
    double* part = nullptr;        // per-step, per-tile partial energies ([chunk_cap][ntl][2])
    int64_t part_cap = 0;          // steps the partial buffer holds
    double* energy = nullptr;      // per-step KE, EE (chunk)
    pdm::StepParam* steps = nullptr;
    int chunk_cap = 0;

    // dynamic clock (mirrors pdSolve::SolveDynamic's loop variables)
    double t = 0, dt_cur = 0;
    int step = 0, last_step = 0;
    double ke = 0, ee = 0;
    double last_kernel_ms = 0;
    int64_t last_kernel_launches = 0;

    // fixed Gaussian integration (fgi_on)
    int fgi_n = 0, fgi_ng = 0;
    double* fgi_tab = nullptr;
    pdm::d4* gauss = nullptr;          // [N][ng] Gauss-point displacements
    pdm::d4* fgi_trial = nullptr;      // [2][N] RK4 trial displacements the integration points interpolate
    pdm::d4* fgi_u0 = nullptr;         // [N] U at the start of the step
    int64_t* cb_ptr = nullptr;         // MLS contributing nodes (n > 1)
    int32_t* cb = nullptr;
    double* phi = nullptr;
    double* tc64 = nullptr;            // tangent weights coef / r^3 per internal entry (static solve)
    float* tc32 = nullptr;             // the same rounded to FP32 (FP32 precisions' CG matvecs)
    int16_t* tc16 = nullptr;           // PDM_FP32C2: the 16 significand bits of tc64 below tc32
                                       // (then tc32 is tc64 rounded toward zero)
    int64_t* fgi_pair_ptr = nullptr;   // [N + 1] live (entry, Gauss point) pairs per internal row
    uint2* fgi_pairs = nullptr;        // (entry offset << 16 | point, j) per live pair
    int64_t fgi_npairs = 0;
    int64_t ncb = 0;                   // MLS contributing-list entries (cb)

    // adaptive integration (ai_on): per internal row, the trapezoidal points of its entries
    int ai = 0;                        // 1 with MLS, 2 without
    int ai_nseg = 0;                   // trapezoidal segments per direction, 2^(intg_point - 1)
    int64_t ai_np = 0, ai_nphi = 0;    // points, MLS coefficients
    int64_t* ai_ptr = nullptr;         // [N + 1]
    pdm::d4* ai_xw = nullptr;          // [P] x, y, z relative to node i, weight
    uint32_t* ai_eo = nullptr;         // [P] entry offset in the internal row
    int64_t* ai_po = nullptr;          // [P] first coefficient in ai_phi
    int32_t* ai_j = nullptr;           // [P] nodes j and i (internal ids) of the point's entry
    int32_t* ai_i = nullptr;
    int32_t* ai_ps = nullptr;          // [P] coefficient stride (points of the entry)
    pdm::d4* ai_res = nullptr;         // [N] family sums (fx, fy, fz, energy density) of a stage
    double* ai_phi = nullptr;

    // graph cache: the captured chunks of RK4 steps, keyed by their step count (a run of
    // nsteps uses at most two: full chunks and the remainder)
    static constexpr int kGraphSlots = 2;
    hipGraphExec_t graph_exec[kGraphSlots] = {nullptr, nullptr};
    int graph_steps[kGraphSlots] = {0, 0};
    int graph_next = 0;             // slot the next capture replaces
};

namespace pdm {

// host model setup (pdm_model.cpp)
int build_nodes(pdm_handle* h);
void init_materials(pdm_handle* h);
int count_max_family(double horizon, double dx, double tol);
double bdryeff_type_one(const Deck& d, double spring, double E, double h);
double time_factor(const BoundaryCondition& b, double t);
bool is_bc_node(const Deck& d, int bcid);
void fill_spec_tables(const pdm_handle* h, SpecDev& sp, MatDev* mats, BCDev* bcs);

// reference-order family search and boundary-effect factors (pdm_bonds.hip)
int build_bonds(pdm_handle* h);
int scan_exclusive(pdm_handle* h, const int64_t* in, int64_t* out, int64_t n);

// tiling, internal CSR, halos, coefficients (pdm_tiles.hip)
int build_tiles(pdm_handle* h);
TileView tile_view(const pdm_handle* h);

// kernels (pdm_kernels.hip)
hipError_t launch_stage(int stage, int precision, const StageArgs& a, int step_idx, hipStream_t s);
hipError_t copy_add_i64(int64_t* dst, const int64_t* src, int64_t n, int64_t off, hipStream_t s);
hipError_t copy_add_i32(int32_t* dst, const int32_t* src, int64_t n, int32_t off, hipStream_t s);
hipError_t fill_i32(int32_t* dst, int64_t n, int32_t v, hipStream_t s);
hipError_t launch_energy_steps(const double* part, int64_t ntl, const int64_t* spec_tile_ptr, int nspec, int nsteps,
                               double* out, hipStream_t s);
hipError_t launch_forces(int precision, const TileView& tv, const d4* u, d4* F, double* elast, hipStream_t s);
// y = K u; with dot set, also each tile's share of u . y at dot[2 k] (k the launch index,
// dot[2 k + 1] = 0), in the layout of the CG reductions' partial pairs.
hipError_t launch_tangent(int precision, const TileView& tv, const d4* u, d4* y, hipStream_t s,
                          double* dot = nullptr);
// Diagonal of each node's 3x3 tangent block into y; with off set, its off-diagonal entries
// (12, 13, 23) into off.
hipError_t launch_tangent_diag(const TileView& tv, const d4* u, d4* y, hipStream_t s, d4* off = nullptr);
// Tangent weights coef / r^3 per internal entry in FP64 (tc64) and rounded to FP32 (tc32, may
// be null); the tangent operations stream them through the view's coef64 / coef32.
hipError_t launch_tangent_coef(const TileView& tv, double* tc64, float* tc32, int16_t* tc16, hipStream_t s);
// The tile view of the tangent operator: coef64 / coef32 replaced by the tangent weights,
// built on first use and kept with the tile layout.
int tangent_view(pdm_handle* h, TileView& tv);
// Per-entry beta corrections of the FP32 force kernels (from coef64); *clamped counts the
// entries whose correction left the 8-bit range.
hipError_t launch_fly_corr(const TileView& tv, const d4* u, uint8_t* corr, int* clamped, hipStream_t s);

// fixed Gaussian integration (pdm_fgi.hip, kernels in pdm_kernels.hip)
int build_fgi(pdm_handle* h);
void free_fgi(pdm_handle* h);
// adaptive integration (pdm_fgi.hip): the points and MLS coefficients; stages run through the
// fixed-Gaussian stage kernels with FgiArgs::ai set
int build_ai(pdm_handle* h);
void free_ai(pdm_handle* h);
// Adaptive-integration family sums of every node into out (fx, fy, fz, energy density): node i's
// displacement from uin, the points' from the MLS interpolation of src (the stage's trial); mode
// 0 prescribes BC nodes' points at the time of step step_idx, mode 2 interpolates every point.
hipError_t launch_ai_family(const StageArgs& a, const FgiArgs& f, const d4* uin, const d4* src, int mode, int step_idx,
                            d4* out, hipStream_t s);
FgiArgs fgi_args(const pdm_handle* h);
hipError_t launch_fgi_gauss(const FgiArgs& f, const d4* u, int step_idx, hipStream_t s);
hipError_t launch_fgi_stage(int stage, const StageArgs& a, const FgiArgs& f, int step_idx, hipStream_t s);
hipError_t launch_fgi_forces(const StageArgs& a, const FgiArgs& f, const d4* u, d4* F, double* elast, hipStream_t s);
// Live-pair lists: fill = false counts each row's live pairs (cnt) and the entries that hold
// one (ent); fill = true writes them at f.pair_ptr.
hipError_t launch_fgi_pairs(const TileView& tv, const FgiArgs& f, bool fill, int64_t* cnt, int64_t* ent,
                            uint2* pairs, hipStream_t s);

// force sums with no allocation in the call (pdm_noalloc.hip): the per-device scratch
// address is looked up once per device at set-up; noalloc_forces enqueues on h->stream
hipError_t noalloc_scratch_setup(int dev);
hipError_t noalloc_forces(const pdm_handle* h, const TileView& tv, const double* u, double* F, double* elast);

// areal force density: sigma[9 t ..] of source node src[t], collinear ids [t][3][ns]
hipError_t launch_afd(const AfdArgs& a, const int64_t* src, int64_t nsrc, const int32_t* colinear, double* sigma,
                      hipStream_t s);

// static solve (pdm_cg.cpp / pdm_cg.hip)
int solve_static(pdm_handle* h, double rel_tol, int max_iterations, pdm_static_stats* stats);
size_t stage_lds_bytes(const TileView& tv, bool with_ecrit, bool with_epi, bool f32);
int stage_vgpr_cap(const TileView& tv, int precision, size_t lds);

}  // namespace pdm
