// pdm_api.cpp -- the C ABI (include/pdm.h): handle lifecycle, InitialProblem on the GPU,
// the RK4 driver with hipGraph-captured steps, state transfer and output writers.
#include <algorithm>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <unordered_map>

#include "pdm_internal.h"

using namespace pdm;

namespace {

thread_local std::string g_create_err;

#define HCHK(x)                                                                                        \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) {                                                                        \
            const_cast<pdm_handle*>(h)->err = std::string("HIP error ") + hipGetErrorString(e_) + " at " #x; \
            return PDM_E_HIP;                                                                          \
        }                                                                                              \
    } while (0)

int fail(const pdm_handle* h, int code, const std::string& m)
{
    const_cast<pdm_handle*>(h)->err = m;
    return code;
}

bool read_file(const char* path, std::string& out)
{
    std::ifstream in(path, std::ios::in | std::ios::binary);
    if (!in) return false;
    std::ostringstream ss;
    ss << in.rdbuf();
    out = ss.str();
    return true;
}

int create_common(const std::string& text, const pdm_options* opt, pdm_handle** out)
{
    if (!out) {
        g_create_err = "null output handle pointer";
        return PDM_E_ARG;
    }
    *out = nullptr;
    pdm_handle* h = new pdm_handle();
    if (opt) h->opt = *opt;
    else pdm_options_default(&h->opt);
    if (h->opt.precision != PDM_FP64 && h->opt.precision != PDM_FP32C && h->opt.precision != PDM_FP32C2) {
        g_create_err = "unknown precision";
        delete h;
        return PDM_E_ARG;
    }
    if (h->opt.graph_chunk <= 0) h->opt.graph_chunk = 50;
    std::string err;
    int rc = parse_deck(text, h->deck, err);
    if (rc) {
        g_create_err = err;
        delete h;
        return rc;
    }
    const Deck& d = h->deck;
    if (!d.cci_on && !d.fgi_on) {
        // adaptive integration: trapezoidal index 1..5 (2 to 17 points per direction), no
        // error-controlled refinement (its point counts depend on the run-time displacement)
        if (d.err_on) {
            g_create_err = "adaptive integration with error_control_on 1 is not built (run-time point refinement)";
            delete h;
            return PDM_E_MODEL;
        }
        if (d.intg_point < 1 || d.intg_point > 5) {
            g_create_err = "intg_point must be 1..5 for adaptive integration (2^(n-1)+1 points per direction)";
            delete h;
            return PDM_E_PARSE;
        }
    } else if (!d.cci_on) {
        const int n = d.intg_point;
        if (n < 1 || n == 7 || n > 8) {
            g_create_err = "intg_point must be 1..6 or 8 (the reference's Gauss rules)";
            delete h;
            return PDM_E_PARSE;
        }
    }
    if (!d.cci_on) {
        if (d.mls_on && !(d.rw > 0.0)) {
            g_create_err = "support_domain_size must be positive";
            delete h;
            return PDM_E_PARSE;
        }
        if (h->opt.precision != PDM_FP64 && !d.fgi_on) {
            g_create_err = "adaptive integration runs in PDM_FP64 only";
            delete h;
            return PDM_E_MODEL;
        }
    }
    if (d.mats.size() > (size_t)kMaxMat || d.bcs.size() > (size_t)kMaxBC || d.slits.size() > (size_t)kMaxSlit) {
        g_create_err = "too many materials / boundary conditions / slits";
        delete h;
        return PDM_E_LIMIT;
    }
    for (const auto& m : d.mats)
        if (m.type != 1) {
            g_create_err = "state-based materials are not supported (the reference's state-based path is commented out)";
            delete h;
            return PDM_E_MODEL;
        }
    *out = h;
    return PDM_OK;
}

template <class T>
void dfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

int require_init(const pdm_handle* h)
{
    if (!h) return PDM_E_ARG;
    if (!h->inited) return fail(h, PDM_E_STATE, "pdm_init has not completed on this handle");
    return PDM_OK;
}

StageArgs stage_args(pdm_handle* h)
{
    StageArgs a{};
    a.tv = tile_view(h);
    a.mat = h->mati;
    a.bc = h->bci;
    a.nb = h->nb;
    a.U = h->U;
    a.V = h->V;
    a.Vb = h->Vb;
    a.k1 = h->k1;
    a.k23 = h->k23;
    a.elast = h->elast;
    a.part = h->part;
    a.steps = h->steps;
    a.mats = h->mats_d;
    a.bcs = h->bcs_d;
    a.nspec = h->nspec;
    a.energies = h->opt.energies ? 1 : 0;
    return a;
}

// Per-entry beta corrections of the FP32 force kernels for the current layout and
// specimen constants (the FP32 records depend on both).
int build_corr(pdm_handle* h)
{
    // fixed Gaussian integration forms its FP32 bonds from the FP64 tables: no corrections
    if (h->opt.precision == PDM_FP64 || h->ntl == 0 || !h->deck.cci_on) return PDM_OK;
    int* cl = nullptr;
    DevTemps temps(h->stream);
    temps.add(&cl);
    HCHK(hipMalloc(&cl, sizeof(int)));
    HCHK(hipMemsetAsync(cl, 0, sizeof(int), h->stream));
    HCHK(launch_fly_corr(tile_view(h), h->u[0], h->corr, cl, h->stream));
    int clamped = 0;
    HCHK(hipStreamSynchronize(h->stream));
    HCHK(hipMemcpy(&clamped, cl, sizeof(int), hipMemcpyDeviceToHost));
    if (h->opt.verbose)
        fprintf(stderr, "pdm_init: %d family entries outside the 8-bit coefficient correction range\n", clamped);
    // a clamped correction leaves that entry's FP32 coefficient off by more than 2^-27
    // relative: refuse the layout rather than run with it
    if (clamped)
        return fail(h, PDM_E_LIMIT, std::to_string(clamped) +
                                        " family entries need an FP32 coefficient correction outside the 8-bit range");
    return PDM_OK;
}

// Constants and tables of a single specimen (after each tile build: the FP32 record
// scale depends on the tile edge).
int upload_spec(pdm_handle* h)
{
    if (!h->specs_d) {
        HCHK(hipMalloc(&h->specs_d, sizeof(SpecDev)));
        HCHK(hipMalloc(&h->mats_d, sizeof(MatDev) * kMaxMat));
        HCHK(hipMalloc(&h->bcs_d, sizeof(BCDev) * kMaxBC));
        HCHK(hipMalloc(&h->spec_tile_ptr_d, sizeof(int64_t) * 2));
    }
    SpecDev sp{};
    std::vector<MatDev> mats(kMaxMat);
    std::vector<BCDev> bcs(kMaxBC);
    fill_spec_tables(h, sp, mats.data(), bcs.data());
    const int64_t tp[2] = {0, h->ntl};
    HCHK(hipMemcpy(h->specs_d, &sp, sizeof sp, hipMemcpyHostToDevice));
    HCHK(hipMemcpy(h->mats_d, mats.data(), sizeof(MatDev) * kMaxMat, hipMemcpyHostToDevice));
    HCHK(hipMemcpy(h->bcs_d, bcs.data(), sizeof(BCDev) * kMaxBC, hipMemcpyHostToDevice));
    HCHK(hipMemcpy(h->spec_tile_ptr_d, tp, sizeof tp, hipMemcpyHostToDevice));
    return build_corr(h);
}

// Host AoS (reference order) <-> device d4 (internal order).
void to_internal(const pdm_handle* h, const double* src, std::vector<d4>& dst)
{
    int64_t n = h->n;
    dst.resize(n);
    for (int64_t a = 0; a < n; ++a) {
        int64_t i = h->perm[a];
        dst[a] = d4{src[3 * i], src[3 * i + 1], src[3 * i + 2], 0.0};
    }
}

void from_internal(const pdm_handle* h, const std::vector<d4>& src, double* dst)
{
    int64_t n = h->n;
    for (int64_t a = 0; a < n; ++a) {
        int64_t i = h->perm[a];
        dst[3 * i] = src[a].x;
        dst[3 * i + 1] = src[a].y;
        dst[3 * i + 2] = src[a].z;
    }
}

// Upload state U, V (3N AoS, reference order); the next stage 1 gathers U.
int upload_state(pdm_handle* h, const double* U, const double* V)
{
    int64_t n = h->n;
    std::vector<d4> buf;
    if (U) {
        to_internal(h, U, buf);
        HCHK(hipMemcpy(h->U, buf.data(), sizeof(d4) * n, hipMemcpyHostToDevice));
        HCHK(hipMemcpy(h->u[0], buf.data(), sizeof(d4) * n, hipMemcpyHostToDevice));
        if (h->fgi_ng) {  // the non-BC nodes' Gauss points follow U, as at the end of a step
            FgiArgs f = fgi_args(h);
            f.mode = 1;
            HCHK(launch_fgi_gauss(f, h->u[0], 0, h->stream));
            HCHK(hipStreamSynchronize(h->stream));
        }
    }
    if (V) {
        to_internal(h, V, buf);
        HCHK(hipMemcpy(h->V, buf.data(), sizeof(d4) * n, hipMemcpyHostToDevice));
    }
    return PDM_OK;
}

int alloc_parts(pdm_handle* h);

// Release every captured chunk of steps.
void drop_graphs(pdm_handle* h)
{
    for (int q = 0; q < pdm_handle::kGraphSlots; ++q) {
        if (h->graph_exec[q]) (void)hipGraphExecDestroy(h->graph_exec[q]);
        h->graph_exec[q] = nullptr;
        h->graph_steps[q] = 0;
    }
    h->graph_next = 0;
}

int ensure_chunk(pdm_handle* h, int cap)
{
    if (h->chunk_cap >= cap) return PDM_OK;
    dfree(h->steps);
    dfree(h->energy);
    HCHK(hipMalloc(&h->steps, sizeof(StepParam) * cap * h->nspec));
    HCHK(hipMalloc(&h->energy, sizeof(double) * 2 * cap * h->nspec));
    h->chunk_cap = cap;
    drop_graphs(h);  // the captured launches point at the old buffers
    return h->tiles ? alloc_parts(h) : PDM_OK;  // one row of tile partials per step
}

// The stage arguments of a batch's tile class: its workgroup map, LDS sizes and layout.
StageArgs class_args(const StageArgs& a, const TileClass& c)
{
    StageArgs b = a;
    b.tv.lmap = c.map;
    b.tv.nmap = c.n;
    b.tv.halo_max = c.halo_max;
    b.tv.tile_max = c.tile_max;
    b.tv.waves = c.waves;
    b.tv.fam_per_wave = c.fams;
    return b;
}

// Enqueue `count` steps whose parameters sit in h->steps[0..count).  A batch with tile
// classes launches each stage once per class (or, with only >= 0, for that class alone,
// which the autotuner times).
int enqueue_steps(pdm_handle* h, int count, int only = -1)
{
    StageArgs a = stage_args(h);
    const bool fgi = !h->deck.cci_on;  // fixed Gaussian or adaptive integration
    const bool gauss = h->fgi_ng > 0;
    const FgiArgs f = fgi ? fgi_args(h) : FgiArgs{};
    std::vector<StageArgs> ca;
    if (!fgi)
        for (size_t c = 0; c < h->tclass.size(); ++c)
            if (only < 0 || (int)c == only) ca.push_back(class_args(a, h->tclass[c]));
    for (int k = 0; k < count; ++k) {
        for (int st = 0; st < 4; ++st) {
            a.uin = h->u[st & 1];
            a.uout = h->u[(st + 1) & 1];
            if (!fgi) {
                if (h->tclass.empty()) {
                    HCHK(launch_stage(st, h->opt.precision, a, k, h->stream));
                    continue;
                }
                for (StageArgs& b : ca) {
                    b.uin = a.uin;
                    b.uout = a.uout;
                    HCHK(launch_stage(st, h->opt.precision, b, k, h->stream));
                }
                continue;
            }
            // Gauss points follow each stage's new displacements: stages 2-4 read them, and
            // the next step's stage 1 reads those of the updated U (pdDataManager.cpp:2043-2125)
            // adaptive integration: the points interpolate this stage's trial displacements
            // (displStep: U at stage 1, then the trial the previous stage wrote)
            if (f.ai)
                HCHK(launch_ai_family(a, f, a.uin, st == 0 ? a.uin : f.trial + ((st + 1) & 1) * f.n, 0, k, h->ai_res,
                                      h->stream));
            HCHK(launch_fgi_stage(st, a, f, k, h->stream));
            if (gauss) HCHK(launch_fgi_gauss(f, st == 3 ? a.uout : f.trial + (st & 1) * f.n, k, h->stream));
        }
    }
    // per-step KE / EE totals, one launch for the chunk
    if (a.energies)
        HCHK(launch_energy_steps(h->part, h->ntl, h->spec_tile_ptr_d, h->nspec, count, h->energy, h->stream));
    return PDM_OK;
}

// The graph of a `count`-step chunk: from the cache (one per distinct step count, kGraphSlots
// of them), else captured from enqueue_steps and instantiated.  The chunk buffers must already
// hold `count` steps.  A captured graph reads its step parameters from h->steps at replay.
int chunk_graph(pdm_handle* h, int count, hipGraphExec_t* out)
{
    *out = nullptr;
    for (int q = 0; q < pdm_handle::kGraphSlots; ++q)
        if (h->graph_exec[q] && h->graph_steps[q] == count) *out = h->graph_exec[q];
    if (*out) return PDM_OK;
    const int q = h->graph_next;
    h->graph_next = (q + 1) % pdm_handle::kGraphSlots;
    if (h->graph_exec[q]) (void)hipGraphExecDestroy(h->graph_exec[q]);
    h->graph_exec[q] = nullptr;
    h->graph_steps[q] = 0;
    hipGraph_t g;
    HCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_steps(h, count);
    hipError_t ce = hipStreamEndCapture(h->stream, &g);
    if (rc) return rc;
    HCHK(ce);
    hipError_t ie = hipGraphInstantiate(&h->graph_exec[q], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HCHK(ie);
    h->graph_steps[q] = count;
    *out = h->graph_exec[q];
    return PDM_OK;
}

// Run `count` steps; sp holds count rows of nspec StepParams ([step][specimen]).
int run_chunk(pdm_handle* h, const std::vector<StepParam>& sp)
{
    const int count = (int)(sp.size() / (size_t)h->nspec);
    if (count == 0) return PDM_OK;
    int rc = ensure_chunk(h, std::max(count, h->opt.graph_chunk));
    if (rc) return rc;
    HCHK(hipMemcpyAsync(h->steps, sp.data(), sizeof(StepParam) * sp.size(), hipMemcpyHostToDevice, h->stream));
    // every chunk runs as a captured graph
    hipGraphExec_t exec = nullptr;
    if (h->opt.use_graph) {
        rc = chunk_graph(h, count, &exec);
        if (rc) return rc;
    }
    HCHK(hipEventRecord(h->ev0, h->stream));
    if (exec) HCHK(hipGraphLaunch(exec, h->stream));
    else {
        rc = enqueue_steps(h, count);
        if (rc) return rc;
    }
    HCHK(hipEventRecord(h->ev1, h->stream));
    HCHK(hipEventSynchronize(h->ev1));
    float ms = 0.f;
    HCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->last_kernel_ms += ms;
    h->last_kernel_launches += 4LL * count;
    if (h->opt.energies) {
        h->last_energy.resize(2 * (size_t)h->nspec);
        HCHK(hipMemcpy(h->last_energy.data(), h->energy + 2 * (int64_t)(count - 1) * h->nspec,
                       sizeof(double) * 2 * h->nspec, hipMemcpyDeviceToHost));
        h->ke = h->last_energy[0];
        h->ee = h->last_energy[1];
    }
    return PDM_OK;
}

// The SolveDynamic loop bookkeeping (pdSolve.cpp:55-120) for up to `want` steps.
void plan_steps(pdm_handle* h, int want, std::vector<StepParam>& sp)
{
    const Deck& d = h->deck;
    sp.clear();
    for (int k = 0; k < want; ++k) {
        int iter = h->step + 1;
        if (iter > d.total_step) break;
        double t = h->t + h->dt_cur;
        if (!(t <= d.total_time)) break;
        h->t = t;
        sp.push_back(StepParam{t, h->dt_cur});
        h->step = iter;
        if (h->t + h->dt_stable > d.total_time && h->t != d.total_time) {
            h->last_step = iter + 1;
            h->dt_cur = d.total_time - h->t;
        }
    }
}

constexpr int64_t kTuneMinEntries = 4000000;  // smaller specimens take the default layout

// Release the tile layout (everything build_tiles allocates), so it can be rebuilt.
void free_classes(pdm_handle* h)
{
    for (TileClass& c : h->tclass) dfree(c.map);
    h->tclass.clear();
    h->spec_class.clear();
}

void free_tiles(pdm_handle* h)
{
    free_classes(h);
    dfree(h->permd); dfree(h->Xi); dfree(h->u[0]); dfree(h->u[1]); dfree(h->rpi); dfree(h->slot);
    dfree(h->self_slot); dfree(h->coef64); dfree(h->coef32); dfree(h->bk); dfree(h->corr); dfree(h->mati);
    dfree(h->bci); dfree(h->nb); dfree(h->tiles); dfree(h->tref); dfree(h->tile_start); dfree(h->halo_ptr);
    dfree(h->halo); dfree(h->part); dfree(h->tc64); dfree(h->tc32); dfree(h->tc16);
    h->part_cap = 0;
    h->ntl = 0;
    h->ntiles = 0;
}

// Per-step, per-tile energy partials for the current layout and chunk capacity.
int alloc_parts(pdm_handle* h)
{
    dfree(h->part);
    const int64_t steps = std::max<int64_t>(h->chunk_cap, 1);
    const size_t bytes = sizeof(double) * 2 * (size_t)std::max<int64_t>(h->ntl, 1) * (size_t)steps;
    HCHK(hipMalloc(&h->part, bytes));
    HCHK(hipMemsetAsync(h->part, 0, bytes, h->stream));
    h->part_cap = steps;
    return PDM_OK;
}

// Zero the dynamic state and the stage ping-pong buffers.
int zero_state(pdm_handle* h)
{
    const int64_t n = std::max<int64_t>(h->n, 1);
    for (d4* p : {h->U, h->V, h->Vb, h->k1, h->k23, h->u[0], h->u[1]})
        HCHK(hipMemsetAsync(p, 0, sizeof(d4) * n, h->stream));
    HCHK(hipMemsetAsync(h->elast, 0, sizeof(double) * n, h->stream));
    if (h->gauss) HCHK(hipMemsetAsync(h->gauss, 0, sizeof(d4) * n * h->fgi_ng, h->stream));
    HCHK(hipStreamSynchronize(h->stream));
    return PDM_OK;
}

// Device, stream and events of a handle that does not share its batch's stream.
int init_device(pdm_handle* h)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= h->opt.device || h->opt.device < 0)
        return fail(h, PDM_E_HIP, "no HIP device " + std::to_string(h->opt.device));
    h->dev = h->opt.device;
    HCHK(hipSetDevice(h->dev));
    if (!h->stream) {
        HCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        HCHK(hipEventCreate(&h->ev0));
        HCHK(hipEventCreate(&h->ev1));
    }
    HCHK(noalloc_scratch_setup(h->dev));  // once per device: the no-allocation entry point's scratch
    return PDM_OK;
}

// InitNode, InitMaterial, traction densities, InitBond and bdryEff, the stable time step.
int init_model(pdm_handle* h)
{
    int rc = build_nodes(h);
    if (rc) return rc;
    init_materials(h);
    // traction density of each TRACTION bc (extension): total force / (nodes carrying it) / dx^3
    const Deck& d = h->deck;
    std::memset(h->trac_density, 0, sizeof(h->trac_density));
    if (d.traction_body) {
        std::vector<int64_t> cnt(d.bcs.size(), 0);
        for (int64_t i = 0; i < h->n; ++i)
            if (h->bc[i] >= 0) cnt[h->bc[i]]++;
        double vol = std::pow(d.dx, 3.0);
        for (size_t b = 0; b < d.bcs.size(); ++b)
            if (d.bcs[b].type == 4 && cnt[b] > 0)
                for (int c = 0; c < 3; ++c)
                    if (d.bcs[b].f[c] == 1) h->trac_density[b][c] = d.bcs[b].v[c] / (double)cnt[b] / vol;
    }
    rc = build_bonds(h);
    if (rc) return rc;
    // CalcStableTimeStep (pdDataManager2.cpp:6-10)
    h->dt_stable = d.fixed_dt_on ? d.fixed_dt : d.safety * h->min_bond / h->max_c;
    return PDM_OK;
}

// Tile layout from the options (0 = default or autotuned).
int layout_options(pdm_handle* h)
{
    if (h->opt.tile_edge) {
        if (h->opt.tile_edge < 2 || h->opt.tile_edge > 7) return fail(h, PDM_E_ARG, "tile_edge must be 2..7");
        h->tile_edge[0] = h->tile_edge[1] = h->tile_edge[2] = h->opt.tile_edge;
    }
    if (h->opt.waves_per_tile) {
        if (h->opt.waves_per_tile != 4 && h->opt.waves_per_tile != 8)
            return fail(h, PDM_E_ARG, "waves_per_tile must be 4 or 8");
        h->waves = h->opt.waves_per_tile;
    }
    if (h->opt.families_per_wave) {
        int f = h->opt.families_per_wave;
        if (f != 1 && f != 2 && f != 4) return fail(h, PDM_E_ARG, "families_per_wave must be 1, 2 or 4");
        h->fam_per_wave = f;
    }
    return PDM_OK;
}

// FP64 dynamic state arrays (N nodes), zeroed.
int alloc_state(pdm_handle* h)
{
    int64_t n = std::max<int64_t>(h->n, 1);
    for (d4** p : {&h->U, &h->V, &h->Vb, &h->k1, &h->k23}) {
        HCHK(hipMalloc(p, sizeof(d4) * n));
        HCHK(hipMemsetAsync(*p, 0, sizeof(d4) * n, h->stream));
    }
    HCHK(hipMalloc(&h->elast, sizeof(double) * n));
    HCHK(hipMemsetAsync(h->elast, 0, sizeof(double) * n, h->stream));
    return PDM_OK;
}

// The driver's clock at t = 0.
void start_clock(pdm_handle* h)
{
    h->dt_cur = h->dt_stable;
    h->t = 0.0;
    h->step = 0;
    h->last_step = h->deck.total_step;
}

// Autotuner: time whole RK4 steps on this specimen for every candidate tile edge,
// wavefronts per tile and families per wavefront, keep the fastest.  The stage
// kernel's speed depends on the halo-to-tile ratio (tile edge), on the LDS and
// register footprint that sets the tiles resident per CU, and on how many lanes share
// a family; the best point differs between precisions and horizons.
// Steps the autotuners run per candidate layout: a few to warm up (the first launch of an
// instantiation also sets its LDS limit), then enough timed steps that the ranking does not
// hinge on the run-to-run spread of a few launches.  The step-parameter and energy-partial
// buffers hold kTuneSteps steps meanwhile.
constexpr int kTuneWarmSteps = 2;
constexpr int kTuneSteps = 10;

int autotune(pdm_handle* h)
{
    static const int kEdges[] = {4, 5, 6};
    static const int kWaves[] = {8, 4};
    static const int kFams[] = {2, 4};
    double best = 1e300;
    int bt = 0, bw = 8, bf = 2;
    int rc = ensure_chunk(h, kTuneSteps);
    if (rc) return rc;
    for (int t : kEdges) {
        h->tile_edge[0] = h->tile_edge[1] = h->tile_edge[2] = t;
        rc = build_tiles(h);
        if (rc == PDM_E_LIMIT) {  // this edge does not fit the LDS or halo limits
            free_tiles(h);
            h->err.clear();
            continue;
        }
        if (rc) return rc;
        rc = alloc_parts(h);
        if (rc) return rc;
        rc = upload_spec(h);
        if (rc) return rc;
        StepParam sp[kTuneSteps];
        for (StepParam& q : sp) q = {h->dt_stable, h->dt_stable};
        HCHK(hipMemcpyAsync(h->steps, sp, sizeof sp, hipMemcpyHostToDevice, h->stream));
        for (int w : kWaves)
            for (int f : kFams) {
                h->waves = w;
                h->fam_per_wave = f;
                rc = enqueue_steps(h, kTuneWarmSteps);
                if (rc) return rc;
                HCHK(hipEventRecord(h->ev0, h->stream));
                rc = enqueue_steps(h, kTuneSteps);
                if (rc) return rc;
                HCHK(hipEventRecord(h->ev1, h->stream));
                HCHK(hipEventSynchronize(h->ev1));
                float ms = 0.f;
                HCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
                const double per = ms / (4.0 * kTuneSteps);
                if (h->opt.verbose)
                    fprintf(stderr, "pdm_init: tune edge %d waves %d families %d: %.4f ms per stage\n", t, w, f, per);
                if (per < best) {
                    best = per;
                    bt = t;
                    bw = w;
                    bf = f;
                }
            }
        free_tiles(h);
    }
    if (!bt) return fail(h, PDM_E_LIMIT, "no tile edge fits this horizon");
    h->tile_edge[0] = h->tile_edge[1] = h->tile_edge[2] = bt;
    h->waves = bw;
    h->fam_per_wave = bf;
    h->tuned = 1;
    h->tune_ms = best;
    return build_tiles(h);
}

// ------------------------------------------------------------ batches
// A batch is one combined handle whose node, entry, tile and halo arrays are the
// specimens' arrays laid end to end (ids and offsets shifted), with a launch-index ->
// specimen map and per-specimen constants: one launch per RK4 stage covers every
// specimen's tiles (a CSR of specimens over the tile schedule).  Each specimen keeps
// its own deck, nodes, reference CSR and clock in a member handle.

// Lay the members' tile layouts (built with the batch's layout) end to end in B.
int batch_concat(pdm_handle* B)
{
    pdm_handle* const h = B;  // for HCHK
    hipStream_t s = B->stream;
    const int S = (int)B->members.size();
    int64_t N = 0, Ei = 0, E = 0, Elive = 0, ntl = 0, ntiles = 0, halo = 0;
    int hmax = 0, tmax = 0, ue = 1;
    for (pdm_handle* m : B->members) {
        N += m->n;
        Ei += m->Ei;
        Elive += m->Elive;
        E += m->E;
        ntl += m->ntl;
        ntiles += m->ntiles;
        halo += m->halo_total;
        hmax = std::max(hmax, m->halo_max);
        tmax = std::max(tmax, m->tile_max);
        ue = ue && m->uniform_mat;
    }
    if (N > INT32_MAX) return fail(B, PDM_E_LIMIT, "a batch holds more than 2^31 - 1 nodes");
    B->n = N;
    B->E = E;
    B->Ei = Ei;
    B->Elive = Elive;
    B->ntl = ntl;
    B->ntiles = ntiles;
    B->halo_total = halo;
    B->halo_max = hmax;
    B->tile_max = tmax;
    B->uniform_mat = ue;
    const bool f64 = B->opt.precision == PDM_FP64;
    const int64_t n1 = std::max<int64_t>(N, 1), e1 = std::max<int64_t>(Ei, 1), t1 = std::max<int64_t>(ntl, 1);
    HCHK(hipMalloc(&B->Xi, sizeof(d4) * n1));
    for (int b = 0; b < 2; ++b) {
        HCHK(hipMalloc(&B->u[b], sizeof(d4) * n1));
        HCHK(hipMemsetAsync(B->u[b], 0, sizeof(d4) * n1, s));
    }
    HCHK(hipMalloc(&B->mati, sizeof(int32_t) * n1));
    HCHK(hipMalloc(&B->bci, sizeof(int32_t) * n1));
    HCHK(hipMalloc(&B->nb, sizeof(int32_t) * n1));
    HCHK(hipMalloc(&B->self_slot, sizeof(uint16_t) * n1));
    HCHK(hipMalloc(&B->rpi, sizeof(int64_t) * (N + 1)));
    HCHK(hipMalloc(&B->slot, sizeof(uint16_t) * e1));
    if (f64) HCHK(hipMalloc(&B->coef64, sizeof(double) * e1));
    else {  // the FP32 stage kernels read no coef32; coef64 only while the corrections are built
        HCHK(hipMalloc(&B->bk, sizeof(BK) * n1));
        HCHK(hipMalloc(&B->corr, e1));
        HCHK(hipMalloc(&B->coef64, sizeof(double) * e1));
    }
    HCHK(hipMalloc(&B->tiles, sizeof(int32_t) * t1));
    HCHK(hipMalloc(&B->tref, sizeof(d4) * t1));
    HCHK(hipMalloc(&B->tile_start, sizeof(int64_t) * (ntiles + 1)));
    HCHK(hipMalloc(&B->halo_ptr, sizeof(int64_t) * (ntl + 1)));
    HCHK(hipMalloc(&B->halo, sizeof(int32_t) * std::max<int64_t>(halo, 1)));
    HCHK(hipMalloc(&B->tile_spec_d, sizeof(int32_t) * t1));
    HCHK(hipMalloc(&B->spec_tile_ptr_d, sizeof(int64_t) * (S + 1)));
    B->spec_node_ptr.assign(S + 1, 0);
    std::vector<int64_t> stp(S + 1, 0);
    int64_t no = 0, eo = 0, to = 0, tio = 0, ho = 0;
    for (int k = 0; k < S; ++k) {
        pdm_handle* m = B->members[k];
        B->spec_node_ptr[k] = no;
        stp[k] = to;
        HCHK(hipMemcpyAsync(B->Xi + no, m->Xi, sizeof(d4) * m->n, hipMemcpyDeviceToDevice, s));
        HCHK(hipMemcpyAsync(B->mati + no, m->mati, sizeof(int32_t) * m->n, hipMemcpyDeviceToDevice, s));
        HCHK(hipMemcpyAsync(B->bci + no, m->bci, sizeof(int32_t) * m->n, hipMemcpyDeviceToDevice, s));
        HCHK(hipMemcpyAsync(B->nb + no, m->nb, sizeof(int32_t) * m->n, hipMemcpyDeviceToDevice, s));
        HCHK(hipMemcpyAsync(B->self_slot + no, m->self_slot, sizeof(uint16_t) * m->n, hipMemcpyDeviceToDevice, s));
        HCHK(copy_add_i64(B->rpi + no, m->rpi, m->n, eo, s));
        HCHK(hipMemcpyAsync(B->slot + eo, m->slot, sizeof(uint16_t) * m->Ei, hipMemcpyDeviceToDevice, s));
        HCHK(hipMemcpyAsync(B->coef64 + eo, m->coef64, sizeof(double) * m->Ei, hipMemcpyDeviceToDevice, s));
        if (!f64) HCHK(hipMemcpyAsync(B->bk + no, m->bk, sizeof(BK) * m->n, hipMemcpyDeviceToDevice, s));
        HCHK(copy_add_i32(B->tiles + to, m->tiles, m->ntl, (int32_t)tio, s));
        HCHK(hipMemcpyAsync(B->tref + to, m->tref, sizeof(d4) * m->ntl, hipMemcpyDeviceToDevice, s));
        HCHK(copy_add_i64(B->tile_start + tio, m->tile_start, m->ntiles, no, s));
        HCHK(copy_add_i64(B->halo_ptr + to, m->halo_ptr, m->ntl, ho, s));
        HCHK(copy_add_i32(B->halo + ho, m->halo, m->halo_total, (int32_t)no, s));
        HCHK(fill_i32(B->tile_spec_d + to, m->ntl, k, s));
        no += m->n;
        eo += m->Ei;
        to += m->ntl;
        tio += m->ntiles;
        ho += m->halo_total;
    }
    B->spec_node_ptr[S] = no;
    stp[S] = to;
    // tile classes: specimens whose tiles fit three or more per CU (LDS), the rest
    free_classes(B);
    std::vector<int> cls(S);
    int ncls[2] = {0, 0};
    for (int k = 0; k < S; ++k) {
        TileView v{};
        v.halo_max = B->members[k]->halo_max;
        v.tile_max = B->members[k]->tile_max;
        const size_t need = stage_lds_bytes(v, !ue, true, !f64) + sizeof(double) * 2 * 8;
        cls[k] = (B->members[k]->ntl == 0 || kLdsBytes / need >= 3) ? 0 : 1;
        if (B->members[k]->ntl) ncls[cls[k]]++;
    }
    if (ncls[0] && ncls[1]) {
        B->tclass.resize(2);
        B->spec_class = cls;
        for (int c = 0; c < 2; ++c) {
            TileClass& tc = B->tclass[c];
            std::vector<int32_t> map;
            for (int k = 0; k < S; ++k) {
                if (cls[k] != c) continue;
                for (int64_t q = stp[k]; q < stp[k + 1]; ++q) map.push_back((int32_t)q);
                tc.halo_max = std::max(tc.halo_max, B->members[k]->halo_max);
                tc.tile_max = std::max(tc.tile_max, B->members[k]->tile_max);
            }
            tc.n = (int64_t)map.size();
            tc.waves = B->waves;
            tc.fams = B->fam_per_wave;
            HCHK(hipMalloc(&tc.map, sizeof(int32_t) * std::max<int64_t>(tc.n, 1)));
            HCHK(hipMemcpy(tc.map, map.data(), sizeof(int32_t) * tc.n, hipMemcpyHostToDevice));
        }
    }
    HCHK(hipMemcpyAsync(B->rpi + N, &Ei, sizeof(int64_t), hipMemcpyHostToDevice, s));
    HCHK(hipMemcpyAsync(B->tile_start + ntiles, &N, sizeof(int64_t), hipMemcpyHostToDevice, s));
    HCHK(hipMemcpyAsync(B->halo_ptr + ntl, &halo, sizeof(int64_t), hipMemcpyHostToDevice, s));
    HCHK(hipMemcpyAsync(B->spec_tile_ptr_d, stp.data(), sizeof(int64_t) * (S + 1), hipMemcpyHostToDevice, s));
    // per-specimen constants and tables
    std::vector<SpecDev> sp(S);
    std::vector<MatDev> mats((size_t)S * kMaxMat);
    std::vector<BCDev> bcs((size_t)S * kMaxBC);
    for (int k = 0; k < S; ++k)
        fill_spec_tables(B->members[k], sp[k], mats.data() + (size_t)k * kMaxMat, bcs.data() + (size_t)k * kMaxBC);
    if (!B->specs_d) {
        HCHK(hipMalloc(&B->specs_d, sizeof(SpecDev) * S));
        HCHK(hipMalloc(&B->mats_d, sizeof(MatDev) * kMaxMat * S));
        HCHK(hipMalloc(&B->bcs_d, sizeof(BCDev) * kMaxBC * S));
    }
    HCHK(hipMemcpyAsync(B->specs_d, sp.data(), sizeof(SpecDev) * S, hipMemcpyHostToDevice, s));
    HCHK(hipMemcpyAsync(B->mats_d, mats.data(), sizeof(MatDev) * kMaxMat * S, hipMemcpyHostToDevice, s));
    HCHK(hipMemcpyAsync(B->bcs_d, bcs.data(), sizeof(BCDev) * kMaxBC * S, hipMemcpyHostToDevice, s));
    HCHK(hipStreamSynchronize(s));
    const int rc = build_corr(B);
    if (!f64) dfree(B->coef64);
    return rc;
}

// Build every member's tiles with edge t, lay them end to end in B, release the members' copies.
int batch_layout(pdm_handle* B, int t)
{
    free_tiles(B);
    dfree(B->tile_spec_d);
    dfree(B->spec_tile_ptr_d);
    for (size_t k = 0; k < B->members.size(); ++k) {
        pdm_handle* m = B->members[k];
        m->tile_edge[0] = m->tile_edge[1] = m->tile_edge[2] = t;
        m->waves = B->waves;
        m->fam_per_wave = B->fam_per_wave;
        int rc = build_tiles(m);
        if (rc) {
            free_tiles(m);
            return fail(B, rc, "specimen " + std::to_string(k) + ": " + m->err);
        }
    }
    B->tile_edge[0] = B->tile_edge[1] = B->tile_edge[2] = t;
    int rc = batch_concat(B);
    for (pdm_handle* m : B->members) free_tiles(m);
    if (rc) return rc;
    // the stage kernels' LDS footprint (FP64 records are the larger)
    size_t need = stage_lds_bytes(tile_view(B), !B->uniform_mat, true, false);
    if (need > kLdsBytes) return fail(B, PDM_E_LIMIT, "tile LDS footprint exceeds the LDS of a CU");
    return alloc_parts(B);
}

// One row of StepParams (dt per specimen) for timing.
int batch_timing_steps(pdm_handle* B)
{
    pdm_handle* const h = B;  // for HCHK
    std::vector<StepParam> sp(kTuneSteps * (size_t)B->nspec);
    for (int r = 0; r < kTuneSteps; ++r)
        for (int k = 0; k < B->nspec; ++k) sp[(size_t)r * B->nspec + k] = {B->members[k]->dt_stable, B->members[k]->dt_stable};
    HCHK(hipMemcpyAsync(B->steps, sp.data(), sizeof(StepParam) * sp.size(), hipMemcpyHostToDevice, B->stream));
    return PDM_OK;
}

// Autotuner over the batch: tile edges 4 and 5, waves per tile 8 and 4, families per wave 2
// and 4; with tile classes each class's layout is timed on its own launches, and an edge
// scores the sum of its classes' best stage times.
int batch_autotune(pdm_handle* B)
{
    pdm_handle* const h = B;  // for HCHK
    static const int kEdges[] = {4, 5};
    static const int kWaves[] = {8, 4};
    static const int kFams[] = {2, 4};
    double best = 1e300;
    int bt = 0;
    std::vector<int> bw, bf;
    std::vector<double> bms;
    int rc = ensure_chunk(B, kTuneSteps);
    if (rc) return rc;
    for (int t : kEdges) {
        rc = batch_layout(B, t);
        if (rc == PDM_E_LIMIT) {
            B->err.clear();
            continue;
        }
        if (rc) return rc;
        rc = batch_timing_steps(B);
        if (rc) return rc;
        const int nc = std::max<int>(1, (int)B->tclass.size());
        std::vector<int> cw(nc, 8), cf(nc, 2);
        std::vector<double> cms(nc, 1e300);
        double total = 0.0;
        for (int c = 0; c < nc; ++c) {
            const int only = B->tclass.empty() ? -1 : c;
            for (int w : kWaves)
                for (int f : kFams) {
                    if (B->tclass.empty()) {
                        B->waves = w;
                        B->fam_per_wave = f;
                    } else {
                        B->tclass[c].waves = w;
                        B->tclass[c].fams = f;
                    }
                    rc = enqueue_steps(B, kTuneWarmSteps, only);
                    if (rc) return rc;
                    HCHK(hipEventRecord(B->ev0, B->stream));
                    rc = enqueue_steps(B, kTuneSteps, only);
                    if (rc) return rc;
                    HCHK(hipEventRecord(B->ev1, B->stream));
                    HCHK(hipEventSynchronize(B->ev1));
                    float ms = 0.f;
                    HCHK(hipEventElapsedTime(&ms, B->ev0, B->ev1));
                    const double per = ms / (4.0 * kTuneSteps);
                    if (B->opt.verbose)
                        fprintf(stderr, "pdm_batch_init: tune edge %d class %d waves %d families %d: %.4f ms per stage\n",
                                t, only, w, f, per);
                    if (per < cms[c]) {
                        cms[c] = per;
                        cw[c] = w;
                        cf[c] = f;
                    }
                }
            total += cms[c];
        }
        if (total < best) {
            best = total;
            bt = t;
            bw = cw;
            bf = cf;
            bms = cms;
        }
    }
    if (!bt) return fail(B, PDM_E_LIMIT, "no tile edge fits these specimens");
    if (B->tile_edge[0] != bt) {
        rc = batch_layout(B, bt);
        if (rc) return rc;
    }
    if (B->tclass.size() != (bw.size() > 1 ? bw.size() : 0)) return fail(B, PDM_E_STATE, "tile classes changed");
    // the reported layout: that of the class with the most tiles
    size_t big = 0;
    for (size_t c = 0; c < B->tclass.size(); ++c) {
        B->tclass[c].waves = bw[c];
        B->tclass[c].fams = bf[c];
        B->tclass[c].tune_ms = bms[c];
        if (B->tclass[c].n > B->tclass[big].n) big = c;
    }
    B->waves = bw[big];
    B->fam_per_wave = bf[big];
    B->tuned = 1;
    B->tune_ms = best;
    return PDM_OK;
}

}  // namespace

namespace pdm {

int tangent_view(pdm_handle* h, TileView& tv)
{
    tv = tile_view(h);
    if (h->ntl == 0) return PDM_OK;
    const bool f32 = h->opt.precision != PDM_FP64;
    const bool lo = h->opt.precision == PDM_FP32C2;
    if (!h->tc64 || (f32 && !h->tc32) || (lo && !h->tc16)) {
        if (!h->coef64) return fail(h, PDM_E_STATE, "the tangent needs the FP64 force coefficients");
        const int64_t ne = std::max<int64_t>(h->Ei, 1);
        if (!h->tc64) HCHK(hipMalloc(&h->tc64, sizeof(double) * ne));
        if (f32 && !h->tc32) HCHK(hipMalloc(&h->tc32, sizeof(float) * ne));
        if (lo && !h->tc16) HCHK(hipMalloc(&h->tc16, sizeof(int16_t) * ne));
        HCHK(launch_tangent_coef(tv, h->tc64, f32 ? h->tc32 : nullptr, lo ? h->tc16 : nullptr, h->stream));
    }
    tv.coef64 = h->tc64;
    tv.coef32 = h->tc32;
    tv.wlo = h->tc16;
    return PDM_OK;
}

}  // namespace pdm

extern "C" {

int pdm_options_default(pdm_options* opt)
{
    if (!opt) return PDM_E_ARG;
    std::memset(opt, 0, sizeof(*opt));
    opt->precision = PDM_FP64;
    opt->device = 0;
    opt->use_graph = 1;
    opt->graph_chunk = 50;
    opt->energies = 1;
    opt->tile_edge = 0;
    opt->waves_per_tile = 0;
    opt->families_per_wave = 0;
    opt->verbose = 0;
    return PDM_OK;
}

int pdm_create(const char* input_path, const pdm_options* opt, pdm_handle** out)
{
    if (!input_path) {
        g_create_err = "null input path";
        return PDM_E_ARG;
    }
    std::string text;
    if (!read_file(input_path, text)) {
        g_create_err = std::string("SetWkDir: input file not exits! (") + input_path + ")";
        return PDM_E_IO;
    }
    return create_common(text, opt, out);
}

int pdm_create_from_text(const char* deck_text, const pdm_options* opt, pdm_handle** out)
{
    if (!deck_text) {
        g_create_err = "null deck text";
        return PDM_E_ARG;
    }
    return create_common(std::string(deck_text), opt, out);
}

int pdm_destroy(pdm_handle* h)
{
    if (!h) return PDM_E_ARG;
    (void)hipSetDevice(h->dev);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graphs(h);
    dfree(h->Xd); dfree(h->rp); dfree(h->col); dfree(h->cut); dfree(h->be); dfree(h->matd);
    dfree(h->permd); dfree(h->Xi); dfree(h->u[0]); dfree(h->u[1]); dfree(h->rpi); dfree(h->slot);
    dfree(h->self_slot); dfree(h->coef64); dfree(h->coef32); dfree(h->bk); dfree(h->corr); dfree(h->mati);
    dfree(h->bci); dfree(h->nb); dfree(h->tiles); dfree(h->tref); dfree(h->tile_start); dfree(h->halo_ptr);
    dfree(h->halo); dfree(h->U); dfree(h->V); dfree(h->Vb); dfree(h->k1); dfree(h->k23);
    dfree(h->elast); dfree(h->part); dfree(h->energy); dfree(h->steps); dfree(h->tc64); dfree(h->tc32); dfree(h->tc16);
    dfree(h->specs_d); dfree(h->mats_d); dfree(h->bcs_d); dfree(h->tile_spec_d); dfree(h->spec_tile_ptr_d);
    free_classes(h);
    free_fgi(h);
    free_ai(h);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream && !h->shared_stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return PDM_OK;
}

int pdm_get_problem_info(const pdm_handle* h, pdm_problem_info* info)
{
    if (!h || !info) return PDM_E_ARG;
    const Deck& d = h->deck;
    std::memset(info, 0, sizeof(*info));
    info->grid_dimension[0] = d.nx;
    info->grid_dimension[1] = d.ny;
    info->grid_dimension[2] = d.nz;
    info->grid_spacing = d.dx;
    info->horizon = d.horizon;
    info->grid_center[0] = d.cx;
    info->grid_center[1] = d.cy;
    info->grid_center[2] = d.cz;
    info->total_time = d.total_time;
    info->total_step = d.total_step;
    info->step_dump_frequency = d.dump;
    info->safety_factor = d.safety;
    info->fixed_time_step = d.fixed_dt_on ? d.fixed_dt : -1.0;
    info->num_materials = (int)d.mats.size();
    info->num_bcs = (int)d.bcs.size();
    info->num_material_spaces = (int)d.mat_spaces.size();
    info->num_deletion_spaces = (int)d.del_spaces.size();
    info->num_boundary_spaces = (int)d.bc_spaces.size();
    info->num_slits = (int)d.slits.size();
    info->damping_coeff = d.damping;
    info->bdryeff_on = d.bdryeff_on;
    info->cci_on = d.cci_on;
    info->cci_modified = d.cci_mdf;
    info->fgi_on = d.fgi_on;
    info->ai_on = d.ai_on;
    info->mls_on = d.mls_on;
    info->intg_point = d.intg_point;
    info->support_domain_size = d.rw;
    info->tolerance = d.tol;
    info->jitter = d.jitter;
    info->jitter_seed = (long long)d.seed;
    info->traction_body = d.traction_body;
    return PDM_OK;
}

const char* pdm_last_error(const pdm_handle* h)
{
    if (!h) return g_create_err.c_str();
    return h->err.c_str();
}

int pdm_init(pdm_handle* h)
{
    if (!h) return PDM_E_ARG;
    if (h->inited) return fail(h, PDM_E_STATE, "pdm_init called twice");
    int rc = init_device(h);
    if (rc) return rc;
    rc = init_model(h);
    if (rc) return rc;
    rc = layout_options(h);
    if (rc) return rc;
    rc = alloc_state(h);
    if (rc) return rc;
    const bool fgi = !h->deck.cci_on;
    const bool tune = !h->opt.tile_edge && !h->opt.waves_per_tile && !h->opt.families_per_wave &&
                      h->E >= kTuneMinEntries && !fgi;
    rc = tune ? autotune(h) : build_tiles(h);
    if (rc) return rc;
    rc = upload_spec(h);
    if (rc) return rc;
    if (fgi) {
        rc = h->deck.fgi_on ? build_fgi(h) : build_ai(h);
        if (rc) return rc;
    }
    rc = alloc_parts(h);
    if (rc) return rc;
    rc = zero_state(h);
    if (rc) return rc;
    h->last_kernel_ms = 0.0;
    h->last_kernel_launches = 0;
    if (h->opt.verbose)
        fprintf(stderr, "pdm_init: nodes=%lld entries=%lld tiles=%lld (edge %dx%dx%d dx) tile_max=%d halo_max=%d "
                        "halo_mean=%.1f lds=%zu B\n",
                (long long)h->n, (long long)h->E, (long long)h->ntl, h->tile_edge[0], h->tile_edge[1],
                h->tile_edge[2], h->tile_max, h->halo_max, h->ntl ? (double)h->halo_total / (double)h->ntl : 0.0,
                stage_lds_bytes(tile_view(h), !h->uniform_mat, true, h->opt.precision != PDM_FP64));
    start_clock(h);
    h->inited = true;
    return PDM_OK;
}

int64_t pdm_num_nodes(const pdm_handle* h) { return h ? h->n : -1; }
int64_t pdm_num_entries(const pdm_handle* h) { return h ? h->E : -1; }
int64_t pdm_num_bonds(const pdm_handle* h) { return h ? h->E / 2 : -1; }
int64_t pdm_num_live_entries(const pdm_handle* h) { return h ? h->Elive : -1; }
int pdm_max_family(const pdm_handle* h) { return h ? h->max_family : -1; }

int pdm_get_nodes(const pdm_handle* h, double* X, int32_t* mat, int32_t* bc)
{
    int rc = require_init(h);
    if (rc) return rc;
    if (X) std::memcpy(X, h->X.data(), sizeof(double) * 3 * h->n);
    if (mat) std::memcpy(mat, h->mat.data(), sizeof(int32_t) * h->n);
    if (bc) std::memcpy(bc, h->bc.data(), sizeof(int32_t) * h->n);
    return PDM_OK;
}

int pdm_get_csr(const pdm_handle* h, int64_t* row_ptr, int32_t* col, uint8_t* cut)
{
    int rc = require_init(h);
    if (rc) return rc;
    if (row_ptr) HCHK(hipMemcpy(row_ptr, h->rp, sizeof(int64_t) * (h->n + 1), hipMemcpyDeviceToHost));
    if (col && h->E) HCHK(hipMemcpy(col, h->col, sizeof(int32_t) * h->E, hipMemcpyDeviceToHost));
    if (cut && h->E) HCHK(hipMemcpy(cut, h->cut, h->E, hipMemcpyDeviceToHost));
    return PDM_OK;
}

int pdm_get_bdryeff(const pdm_handle* h, double* be)
{
    int rc = require_init(h);
    if (rc) return rc;
    if (!be) return PDM_E_ARG;
    HCHK(hipMemcpy(be, h->be, sizeof(double) * h->n, hipMemcpyDeviceToHost));
    return PDM_OK;
}

int pdm_get_timestep(const pdm_handle* h, double* dt, double* min_bond_length)
{
    int rc = require_init(h);
    if (rc) return rc;
    if (dt) *dt = h->dt_stable;
    if (min_bond_length) *min_bond_length = h->min_bond;
    return PDM_OK;
}

int pdm_compute_forces(pdm_handle* h, const double* u, double* F, double* elast)
{
    int rc = require_init(h);
    if (rc) return rc;
    if (!u || !F) return fail(h, PDM_E_ARG, "null u or F");
    int64_t n = h->n;
    if (n == 0) return PDM_OK;
    std::vector<d4> buf;
    to_internal(h, u, buf);
    d4 *ud = nullptr, *Fd = nullptr;
    double* ed = nullptr;
    HCHK(hipMalloc(&ud, sizeof(d4) * n));
    HCHK(hipMalloc(&Fd, sizeof(d4) * n));
    HCHK(hipMalloc(&ed, sizeof(double) * n));
    HCHK(hipMemcpy(ud, buf.data(), sizeof(d4) * n, hipMemcpyHostToDevice));
    if (h->ai) {
        // every node's points interpolated from u (the stepping state recomputes its own per stage)
        HCHK(launch_ai_family(stage_args(h), fgi_args(h), ud, ud, 2, 0, Fd, h->stream));
        HCHK(hipStreamSynchronize(h->stream));
        std::vector<d4> r(n);
        HCHK(hipMemcpy(r.data(), Fd, sizeof(d4) * n, hipMemcpyDeviceToHost));
        std::vector<double> e(n);
        for (int64_t q = 0; q < n; ++q) e[q] = r[q].w;
        HCHK(hipMemcpy(ed, e.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    } else if (h->fgi_ng) {
        // every node's Gauss points interpolated from u, into scratch (the stepping state keeps its own)
        d4* gd = nullptr;
        HCHK(hipMalloc(&gd, sizeof(d4) * n * h->fgi_ng));
        FgiArgs f = fgi_args(h);
        f.mode = 2;
        f.gout = gd;
        f.gin = gd;
        HCHK(launch_fgi_gauss(f, ud, 0, h->stream));
        HCHK(launch_fgi_forces(stage_args(h), f, ud, Fd, ed, h->stream));
        HCHK(hipStreamSynchronize(h->stream));
        HCHK(hipFree(gd));
    } else {
        HCHK(launch_forces(h->opt.precision, tile_view(h), ud, Fd, ed, h->stream));
    }
    HCHK(hipStreamSynchronize(h->stream));
    HCHK(hipMemcpy(buf.data(), Fd, sizeof(d4) * n, hipMemcpyDeviceToHost));
    from_internal(h, buf, F);
    if (elast) {
        std::vector<double> e(n);
        HCHK(hipMemcpy(e.data(), ed, sizeof(double) * n, hipMemcpyDeviceToHost));
        for (int64_t a = 0; a < n; ++a) elast[h->perm[a]] = e[a];
    }
    HCHK(hipFree(ud));
    HCHK(hipFree(Fd));
    HCHK(hipFree(ed));
    return PDM_OK;
}

int pdm_compute_forces_noalloc(pdm_handle* h, const double* u, double* F, double* elast, int sync)
{
    int rc = require_init(h);
    if (rc) return rc;
    if (!u || !F) return fail(h, PDM_E_ARG, "null u or F");
    if (!h->deck.cci_on) return fail(h, PDM_E_MODEL, "pdm_compute_forces_noalloc takes CCI decks only");
    if (h->n > kNoAllocMaxNodes)
        return fail(h, PDM_E_LIMIT, std::to_string(h->n) + " nodes exceed the " + std::to_string(kNoAllocMaxNodes) +
                                        " the no-allocation scratch holds; use pdm_compute_forces");
    if (h->n == 0) return PDM_OK;
    HCHK(hipSetDevice(h->dev));
    HCHK(noalloc_forces(h, tile_view(h), u, F, elast));
    if (sync) HCHK(hipStreamSynchronize(h->stream));
    return PDM_OK;
}

int pdm_areal_force_density(pdm_handle* h, const int64_t* nodes, int64_t count, const double* bdryeff,
                            double* sigma)
{
    int rc = require_init(h);
    if (rc) return rc;
    if (count < 0 || (count > 0 && (!nodes || !sigma))) return fail(h, PDM_E_ARG, "null nodes or sigma");
    if (!h->deck.cci_on) return fail(h, PDM_E_MODEL, "the areal force density is built for CCI decks only");
    if (count == 0) return PDM_OK;
    const Deck& d = h->deck;
    const int64_t n = h->n;
    for (int64_t t = 0; t < count; ++t)
        if (nodes[t] < 0 || nodes[t] >= n) return fail(h, PDM_E_ARG, "node id out of range");
    const int ns = (int)(d.horizon / d.dx);
    if (ns < 1) return fail(h, PDM_E_MODEL, "the horizon is shorter than one grid spacing");
    // collinear lattice nodes by exact coordinate equality (CalcArealForceDensity,
    // pdDataManager3.cpp:463-481): x_s - dx (1 + k) along each axis
    struct Key {
        uint64_t a, b, c;
        bool operator==(const Key& o) const { return a == o.a && b == o.b && c == o.c; }
    };
    struct KeyHash {
        size_t operator()(const Key& k) const
        {
            uint64_t x = k.a * 0x9E3779B97F4A7C15ULL ^ (k.b + 0x632BE59BD9B4E019ULL + (k.a << 6));
            x ^= k.c * 0xC2B2AE3D27D4EB4FULL + (x >> 29);
            return (size_t)(x ^ (x >> 32));
        }
    };
    auto key = [](double x, double y, double z) {
        Key k;
        std::memcpy(&k.a, &x, 8);
        std::memcpy(&k.b, &y, 8);
        std::memcpy(&k.c, &z, 8);
        return k;
    };
    std::unordered_map<Key, int32_t, KeyHash> where;
    where.reserve((size_t)n * 2);
    for (int64_t i = n - 1; i >= 0; --i) where[key(h->X[3 * i], h->X[3 * i + 1], h->X[3 * i + 2])] = (int32_t)i;
    std::vector<int32_t> colin((size_t)count * 3 * ns, -1);
    for (int64_t t = 0; t < count; ++t) {
        const double* xs = &h->X[3 * nodes[t]];
        for (int q = 0; q < 3; ++q)
            for (int k = 0; k < ns; ++k) {
                double y[3] = {xs[0], xs[1], xs[2]};
                y[q] = xs[q] - d.dx * (1 + k);
                auto it = where.find(key(y[0], y[1], y[2]));
                if (it != where.end()) colin[((size_t)t * 3 + q) * ns + k] = it->second;
            }
    }
    // displacements in reference order (and the caller's boundary-effect factors)
    std::vector<double> U(3 * (size_t)n);
    rc = pdm_get_state(h, U.data(), nullptr);
    if (rc) return rc;
    std::vector<d4> Ud(n);
    for (int64_t i = 0; i < n; ++i) Ud[i] = d4{U[3 * i], U[3 * i + 1], U[3 * i + 2], 0.0};
    d4* Udev = nullptr;
    double *bed = nullptr, *sig = nullptr;
    int64_t* srcd = nullptr;
    int32_t* cold = nullptr;
    DevTemps temps(h->stream);
    temps.add(&Udev);
    temps.add(&bed);
    temps.add(&sig);
    temps.add(&srcd);
    temps.add(&cold);
    HCHK(hipMalloc(&Udev, sizeof(d4) * n));
    HCHK(hipMalloc(&sig, sizeof(double) * 9 * count));
    HCHK(hipMalloc(&srcd, sizeof(int64_t) * count));
    HCHK(hipMalloc(&cold, sizeof(int32_t) * colin.size()));
    HCHK(hipMemcpyAsync(Udev, Ud.data(), sizeof(d4) * n, hipMemcpyHostToDevice, h->stream));
    HCHK(hipMemcpyAsync(srcd, nodes, sizeof(int64_t) * count, hipMemcpyHostToDevice, h->stream));
    HCHK(hipMemcpyAsync(cold, colin.data(), sizeof(int32_t) * colin.size(), hipMemcpyHostToDevice, h->stream));
    if (bdryeff) {
        HCHK(hipMalloc(&bed, sizeof(double) * n));
        HCHK(hipMemcpyAsync(bed, bdryeff, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
    }
    AfdArgs a{};
    a.X = h->Xd;
    a.U = Udev;
    a.rp = h->rp;
    a.col = h->col;
    a.cut = h->cut;
    a.mat = h->matd;
    a.be = bdryeff ? bed : h->be;
    for (size_t m = 0; m < d.mats.size() && m < (size_t)kMaxMat; ++m) {
        a.ecrit[m] = d.mats[m].ecrit;
        a.spring[m] = h->spring[m];
    }
    a.horizon = d.horizon;
    a.dx = d.dx;
    a.vol = std::pow(d.dx, 3.0);
    a.ns = ns;
    a.ds = d.horizon / ns;
    a.original_cci = (d.cci_on && !d.cci_mdf) ? 1 : 0;
    HCHK(launch_afd(a, srcd, count, cold, sig, h->stream));
    HCHK(hipMemcpyAsync(sigma, sig, sizeof(double) * 9 * count, hipMemcpyDeviceToHost, h->stream));
    HCHK(hipStreamSynchronize(h->stream));
    return PDM_OK;
}

int64_t pdm_num_ai_points(const pdm_handle* h) { return h ? h->ai_np : -1; }
int64_t pdm_num_ai_coefficients(const pdm_handle* h) { return h ? h->ai_nphi : -1; }

int pdm_get_ai_points(const pdm_handle* h, int32_t* node, int32_t* nbr, double* xw)
{
    int rc = require_init(h);
    if (rc) return rc;
    if (!h->ai) return fail(h, PDM_E_MODEL, "the deck has no adaptive-integration points");
    const int64_t n = h->n, P = h->ai_np;
    std::vector<int64_t> ptr(n + 1);
    std::vector<int32_t> j(P);
    std::vector<d4> x(P);
    HCHK(hipStreamSynchronize(h->stream));
    HCHK(hipMemcpy(ptr.data(), h->ai_ptr, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost));
    if (P) {
        HCHK(hipMemcpy(j.data(), h->ai_j, sizeof(int32_t) * P, hipMemcpyDeviceToHost));
        HCHK(hipMemcpy(x.data(), h->ai_xw, sizeof(d4) * P, hipMemcpyDeviceToHost));
    }
    for (int64_t a = 0; a < n; ++a)
        for (int64_t q = ptr[a]; q < ptr[a + 1]; ++q) {
            if (node) node[q] = h->perm[a];
            if (nbr) nbr[q] = h->perm[j[q]];
            if (xw) {
                xw[4 * q] = x[q].x;
                xw[4 * q + 1] = x[q].y;
                xw[4 * q + 2] = x[q].z;
                xw[4 * q + 3] = x[q].w;
            }
        }
    return PDM_OK;
}

int pdm_num_gauss_points(const pdm_handle* h) { return h ? h->fgi_ng : -1; }
int64_t pdm_num_live_gauss_pairs(const pdm_handle* h) { return h ? h->fgi_npairs : -1; }

int pdm_get_gauss(const pdm_handle* h, double* Xg, double* Ug)
{
    int rc = require_init(h);
    if (rc) return rc;
    const int ng = h->fgi_ng;
    if (ng == 0) return fail(h, PDM_E_MODEL, "the deck has no Gauss points (CCI)");
    const int64_t n = h->n;
    if (Xg) {
        std::vector<double> tab((size_t)kFgiTab * ng);
        HCHK(hipMemcpy(tab.data(), h->fgi_tab, sizeof(double) * tab.size(), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; ++i)
            for (int g = 0; g < ng; ++g)
                for (int c = 0; c < 3; ++c) Xg[(i * ng + g) * 3 + c] = tab[(size_t)kFgiTab * g + c] + h->X[3 * i + c];
    }
    if (Ug) {
        std::vector<d4> G((size_t)n * ng);
        HCHK(hipStreamSynchronize(h->stream));
        HCHK(hipMemcpy(G.data(), h->gauss, sizeof(d4) * G.size(), hipMemcpyDeviceToHost));
        for (int64_t a = 0; a < n; ++a) {
            const int64_t i = h->perm[a];
            for (int g = 0; g < ng; ++g) {
                const d4& v = G[(size_t)a * ng + g];
                double* o = &Ug[(i * ng + g) * 3];
                o[0] = v.x;
                o[1] = v.y;
                o[2] = v.z;
            }
        }
    }
    return PDM_OK;
}

int pdm_apply_tangent(pdm_handle* h, const double* u, double* y)
{
    int rc = require_init(h);
    if (rc) return rc;
    if (!u || !y) return fail(h, PDM_E_ARG, "null u or y");
    if (!h->deck.cci_on) return fail(h, PDM_E_MODEL, "the tangent operator is built for CCI decks only");
    int64_t n = h->n;
    if (n == 0) return PDM_OK;
    std::vector<d4> buf;
    to_internal(h, u, buf);
    d4 *ud = nullptr, *yd = nullptr;
    HCHK(hipMalloc(&ud, sizeof(d4) * n));
    HCHK(hipMalloc(&yd, sizeof(d4) * n));
    HCHK(hipMemcpy(ud, buf.data(), sizeof(d4) * n, hipMemcpyHostToDevice));
    TileView tv;
    rc = tangent_view(h, tv);
    if (rc) return rc;
    HCHK(launch_tangent(h->opt.precision, tv, ud, yd, h->stream));
    HCHK(hipStreamSynchronize(h->stream));
    HCHK(hipMemcpy(buf.data(), yd, sizeof(d4) * n, hipMemcpyDeviceToHost));
    from_internal(h, buf, y);
    HCHK(hipFree(ud));
    HCHK(hipFree(yd));
    return PDM_OK;
}

int pdm_set_state(pdm_handle* h, const double* U, const double* V)
{
    int rc = require_init(h);
    if (rc) return rc;
    HCHK(hipStreamSynchronize(h->stream));
    return upload_state(h, U, V);
}

int pdm_get_state(const pdm_handle* h, double* U, double* V)
{
    int rc = require_init(h);
    if (rc) return rc;
    int64_t n = h->n;
    std::vector<d4> buf(n);
    HCHK(hipStreamSynchronize(h->stream));
    if (U) {
        HCHK(hipMemcpy(buf.data(), h->U, sizeof(d4) * n, hipMemcpyDeviceToHost));
        from_internal(h, buf, U);
    }
    if (V) {
        HCHK(hipMemcpy(buf.data(), h->V, sizeof(d4) * n, hipMemcpyDeviceToHost));
        from_internal(h, buf, V);
    }
    return PDM_OK;
}

int pdm_step_rk4(pdm_handle* h, int nsteps, int* done)
{
    int rc = require_init(h);
    if (rc) return rc;
    if (nsteps < 0) return fail(h, PDM_E_ARG, "negative step count");
    HCHK(hipSetDevice(h->dev));
    h->last_kernel_ms = 0.0;
    h->last_kernel_launches = 0;
    int total = 0;
    std::vector<StepParam> sp;
    while (total < nsteps) {
        int want = std::min(nsteps - total, h->opt.graph_chunk);
        plan_steps(h, want, sp);
        if (sp.empty()) break;
        rc = run_chunk(h, sp);
        if (rc) return rc;
        total += (int)sp.size();
        if ((int)sp.size() < want) break;
    }
    if (done) *done = total;
    return PDM_OK;
}

int pdm_prepare_steps(pdm_handle* h, int steps)
{
    int rc = require_init(h);
    if (rc) return rc;
    if (steps <= 0) return fail(h, PDM_E_ARG, "step count must be positive");
    if (!h->opt.use_graph) return PDM_OK;
    HCHK(hipSetDevice(h->dev));
    rc = ensure_chunk(h, std::max(steps, h->opt.graph_chunk));
    if (rc) return rc;
    hipGraphExec_t exec = nullptr;
    return chunk_graph(h, steps, &exec);
}

int pdm_get_clock(const pdm_handle* h, double* t, int* step, double* dt_next)
{
    int rc = require_init(h);
    if (rc) return rc;
    if (t) *t = h->t;
    if (step) *step = h->step;
    if (dt_next) *dt_next = h->dt_cur;
    return PDM_OK;
}

int pdm_get_energy(const pdm_handle* h, double* ke, double* ee)
{
    int rc = require_init(h);
    if (rc) return rc;
    if (ke) *ke = h->ke;
    if (ee) *ee = h->ee;
    return PDM_OK;
}

int pdm_get_tuning(const pdm_handle* h, int* tile_edge, int* waves_per_tile, int* families_per_wave,
                   double* stage_ms)
{
    int rc = require_init(h);
    if (rc) return rc;
    if (tile_edge) *tile_edge = h->tile_edge[0];
    if (waves_per_tile) *waves_per_tile = h->waves;
    if (families_per_wave) *families_per_wave = h->fam_per_wave;
    if (stage_ms) *stage_ms = h->tuned ? h->tune_ms : 0.0;
    return PDM_OK;
}

int pdm_get_stage_launch(const pdm_handle* h, int* max_vgprs, int64_t* lds_bytes)
{
    int rc = require_init(h);
    if (rc) return rc;
    if (!h->deck.cci_on) return fail(h, PDM_E_MODEL, "fixed Gaussian and adaptive-integration decks step with their own kernels");
    const size_t lds = stage_lds_bytes(tile_view(h), !h->uniform_mat, true, h->opt.precision != PDM_FP64);
    if (max_vgprs) *max_vgprs = stage_vgpr_cap(tile_view(h), h->opt.precision, lds);
    if (lds_bytes) *lds_bytes = (int64_t)lds;
    return PDM_OK;
}

int pdm_last_kernel_timing(const pdm_handle* h, double* avg_ms, int64_t* launches)
{
    int rc = require_init(h);
    if (rc) return rc;
    if (avg_ms) *avg_ms = h->last_kernel_launches ? h->last_kernel_ms / (double)h->last_kernel_launches : 0.0;
    if (launches) *launches = h->last_kernel_launches;
    return PDM_OK;
}

int pdm_write_tecplot(const pdm_handle* h, const char* path, int step)
{
    int rc = require_init(h);
    if (rc) return rc;
    (void)step;
    if (!path) return fail(h, PDM_E_ARG, "null path");
    int64_t n = h->n;
    std::vector<double> U(3 * n), V(3 * n), be(n);
    rc = pdm_get_state(h, U.data(), V.data());
    if (rc) return rc;
    rc = pdm_get_bdryeff(h, be.data());
    if (rc) return rc;
    std::ofstream out(path, std::ios::out);
    if (!out) return fail(h, PDM_E_IO, std::string("cannot write ") + path);
    const Deck& d = h->deck;
    // pdSolve::WriteTecplot (pdSolve.cpp:708-732), pdNode::Print (pdNode.cpp:418-433)
    out << "Tecplot formatting output mode\n";
    out.setf(std::ios::scientific);
    out.precision(5);
    out << "VARIABLES = \n ID X1 X2 X3 U1 U2 U3 V1 V2 V3 MatID BCID bdryEff DmgFrac \n"
        << "ZONE I=" << d.nx << ", J=" << d.ny << ", K=" << d.nz << ", F=POINT\n\n";
    for (int64_t i = 0; i < n; ++i) {
        out << i << "\t" << h->X[3 * i] << "\t" << h->X[3 * i + 1] << "\t" << h->X[3 * i + 2] << "\t" << U[3 * i]
            << "\t" << U[3 * i + 1] << "\t" << U[3 * i + 2] << "\t" << V[3 * i] << "\t" << V[3 * i + 1] << "\t"
            << V[3 * i + 2] << "\t" << h->mat[i];
        out << "\t" << h->bc[i];
        out << "\t" << be[i] << "\t" << 0.0 << std::endl;
    }
    return out.good() ? PDM_OK : fail(h, PDM_E_IO, std::string("write failed: ") + path);
}

int pdm_solve_dynamic(pdm_handle* h, const char* out_dir)
{
    int rc = require_init(h);
    if (rc) return rc;
    std::string dir = out_dir ? std::string(out_dir) : std::string(".");
    if (!dir.empty() && dir.back() != '/') dir += '/';
    std::ofstream outf((dir + "pds_output.dat").c_str(), std::ios::app);
    if (!outf) return fail(h, PDM_E_IO, "cannot open pds_output.dat in " + dir);
    outf.setf(std::ios::scientific);
    outf.precision(5);
    std::ofstream stepf((dir + "pds_steps.dat").c_str(), std::ios::out);
    stepf << "Data at given node for each time step output mode\n\n";
    stepf << "Step Time ID X1 X2 X3 U1 U2 U3 V1 V2 V3 MatID BCID bdryEff DmgFrac \n";
    outf << std::endl << "** Dynamic solution: starting from cycle 1 **" << std::endl;
    const Deck& d = h->deck;
    int dump = std::max(1, d.dump);
    // WriteEachStep (pdSolve.cpp:751-774): every step, the nodes at exactly (dx/2, dx/2, dx/2),
    // printed as pdNode::Print (pdNode.cpp:418-433) after the step and the time, in the step
    // file's default number format
    struct Watch {
        int64_t ref, in;
    };
    std::vector<Watch> watch;
    const double half = 0.5 * d.dx;
    for (int64_t i = 0; i < h->n; ++i)
        if (h->X[3 * i] == half && h->X[3 * i + 1] == half && h->X[3 * i + 2] == half) watch.push_back({i, -1});
    std::vector<double> be;
    if (!watch.empty()) {
        for (int64_t a = 0; a < h->n; ++a)
            for (Watch& w : watch)
                if (h->perm[a] == w.ref) w.in = a;
        be.resize(h->n);
        rc = pdm_get_bdryeff(h, be.data());
        if (rc) return rc;
    }
    std::vector<StepParam> sp;
    std::vector<double> en;
    // break chunks at steps that write a dump (1, every dump, last), and after every step
    // while nodes are watched
    for (;;) {
        int next = h->step + 1;
        int want = 1;
        if (next != 1 && watch.empty()) {
            int to_dump = dump - (h->step % dump);
            want = std::min(to_dump, h->opt.graph_chunk);
        }
        int before = h->step;
        plan_steps(h, want, sp);
        if (sp.empty()) break;
        rc = run_chunk(h, sp);
        if (rc) return rc;
        int cnt = (int)sp.size();
        if (h->opt.energies) {
            en.resize(2 * cnt);
            HCHK(hipMemcpy(en.data(), h->energy, sizeof(double) * 2 * cnt, hipMemcpyDeviceToHost));
        }
        for (int k = 0; k < cnt; ++k) {
            int iter = before + k + 1;
            outf << iter << ": dt= " << sp[k].dt << " t= " << sp[k].t << " KE= " << (h->opt.energies ? en[2 * k] : 0.0)
                 << " EE= " << (h->opt.energies ? en[2 * k + 1] : 0.0) << std::endl;
        }
        for (const Watch& w : watch) {  // one step per chunk here: the state after step h->step
            d4 u, v;
            HCHK(hipMemcpy(&u, h->U + w.in, sizeof(d4), hipMemcpyDeviceToHost));
            HCHK(hipMemcpy(&v, h->V + w.in, sizeof(d4), hipMemcpyDeviceToHost));
            const double* x = &h->X[3 * w.ref];
            stepf << h->step << "\t" << sp[cnt - 1].t << "\t" << w.ref << "\t" << x[0] << "\t" << x[1] << "\t" << x[2]
                  << "\t" << u.x << "\t" << u.y << "\t" << u.z << "\t" << v.x << "\t" << v.y << "\t" << v.z << "\t"
                  << h->mat[w.ref] << "\t" << h->bc[w.ref] << "\t" << be[w.ref] << "\t" << 0.0 << std::endl;
        }
        int iter = h->step;
        if (iter == 1 || iter == h->last_step || iter % dump == 0) {
            std::string path = dir + "pds_tecplot_" + std::to_string(iter) + ".dat";
            rc = pdm_write_tecplot(h, path.c_str(), iter);
            if (rc) return rc;
        }
    }
    return PDM_OK;
}

int pdm_solve_static(pdm_handle* h, double rel_tol, int max_iterations, pdm_static_stats* stats)
{
    int rc = require_init(h);
    if (rc) return rc;
    if (!h->deck.cci_on) return fail(h, PDM_E_MODEL, "the static solve is built for CCI decks only");
    HCHK(hipSetDevice(h->dev));
    return solve_static(h, rel_tol, max_iterations, stats);
}

}  // extern "C"

// ------------------------------------------------------------ batches (C ABI)

struct pdm_batch {
    pdm_handle* h = nullptr;  // combined arrays; the specimens are h->members
};

namespace {

int batch_create(const std::vector<std::string>& texts, const pdm_options* opt, pdm_batch** out)
{
    if (!out) {
        g_create_err = "null output handle pointer";
        return PDM_E_ARG;
    }
    *out = nullptr;
    if (texts.empty()) {
        g_create_err = "a batch needs at least one specimen";
        return PDM_E_ARG;
    }
    pdm_handle* B = nullptr;
    // the combined handle carries the options; its own deck is the first specimen's
    int rc = create_common(texts[0], opt, &B);
    if (rc) {
        g_create_err = "specimen 0: " + g_create_err;
        return rc;
    }
    for (size_t k = 0; k < texts.size(); ++k) {
        pdm_handle* m = nullptr;
        rc = create_common(texts[k], opt, &m);
        if (rc) {
            g_create_err = "specimen " + std::to_string(k) + ": " + g_create_err;
            for (pdm_handle* q : B->members) pdm_destroy(q);
            B->members.clear();
            pdm_destroy(B);
            return rc;
        }
        m->shared_stream = true;
        B->members.push_back(m);
        if (!m->deck.cci_on) {
            g_create_err = "specimen " + std::to_string(k) + ": batches hold CCI decks only";
            for (pdm_handle* q : B->members) pdm_destroy(q);
            B->members.clear();
            pdm_destroy(B);
            return PDM_E_MODEL;
        }
    }
    B->nspec = (int)texts.size();
    pdm_batch* b = new pdm_batch();
    b->h = B;
    *out = b;
    return PDM_OK;
}

int batch_spec(const pdm_batch* b, int spec)
{
    if (!b || !b->h) return PDM_E_ARG;
    if (!b->h->inited) return fail(b->h, PDM_E_STATE, "pdm_batch_init has not run");
    if (spec < 0 || spec >= b->h->nspec) return fail(b->h, PDM_E_ARG, "specimen index out of range");
    return PDM_OK;
}

}  // namespace

extern "C" {

int pdm_batch_create(const char* const* input_paths, int nspec, const pdm_options* opt, pdm_batch** out)
{
    if (!input_paths || nspec <= 0) {
        g_create_err = "no input paths";
        return PDM_E_ARG;
    }
    std::vector<std::string> texts(nspec);
    for (int k = 0; k < nspec; ++k) {
        if (!input_paths[k] || !read_file(input_paths[k], texts[k])) {
            g_create_err = std::string("cannot read input file ") + (input_paths[k] ? input_paths[k] : "(null)");
            return PDM_E_IO;
        }
    }
    return batch_create(texts, opt, out);
}

int pdm_batch_create_from_text(const char* const* deck_texts, int nspec, const pdm_options* opt, pdm_batch** out)
{
    if (!deck_texts || nspec <= 0) {
        g_create_err = "no deck texts";
        return PDM_E_ARG;
    }
    std::vector<std::string> texts(nspec);
    for (int k = 0; k < nspec; ++k) {
        if (!deck_texts[k]) {
            g_create_err = "null deck text";
            return PDM_E_ARG;
        }
        texts[k] = deck_texts[k];
    }
    return batch_create(texts, opt, out);
}

int pdm_batch_destroy(pdm_batch* b)
{
    if (!b) return PDM_E_ARG;
    if (b->h) {
        if (b->h->stream) (void)hipStreamSynchronize(b->h->stream);
        for (pdm_handle* m : b->h->members) pdm_destroy(m);
        b->h->members.clear();
        pdm_destroy(b->h);
    }
    delete b;
    return PDM_OK;
}

const char* pdm_batch_last_error(const pdm_batch* b)
{
    if (!b || !b->h) return g_create_err.c_str();
    return b->h->err.c_str();
}

int pdm_batch_init(pdm_batch* b)
{
    if (!b || !b->h) return PDM_E_ARG;
    pdm_handle* B = b->h;
    if (B->inited) return fail(B, PDM_E_STATE, "pdm_batch_init called twice");
    int rc = init_device(B);
    if (rc) return rc;
    for (size_t k = 0; k < B->members.size(); ++k) {
        pdm_handle* m = B->members[k];
        m->dev = B->dev;
        m->stream = B->stream;
        rc = init_model(m);
        if (rc) return fail(B, rc, "specimen " + std::to_string(k) + ": " + m->err);
        start_clock(m);
        m->inited = true;
    }
    rc = layout_options(B);
    if (rc) return rc;
    int64_t E = 0, N = 0;
    for (pdm_handle* m : B->members) {
        E += m->E;
        N += m->n;
    }
    B->n = N;
    rc = alloc_state(B);
    if (rc) return rc;
    const bool tune = !B->opt.tile_edge && !B->opt.waves_per_tile && !B->opt.families_per_wave &&
                      E >= kTuneMinEntries;
    rc = tune ? batch_autotune(B) : batch_layout(B, B->tile_edge[0]);
    if (rc) return rc;
    rc = zero_state(B);
    if (rc) return rc;
    B->last_kernel_ms = 0.0;
    B->last_kernel_launches = 0;
    if (B->opt.verbose) {
        fprintf(stderr, "pdm_batch_init: %d specimens, nodes=%lld entries=%lld tiles=%lld edge %d waves %d "
                        "families %d halo_max=%d\n",
                B->nspec, (long long)B->n, (long long)B->E, (long long)B->ntl, B->tile_edge[0], B->waves,
                B->fam_per_wave, B->halo_max);
        for (size_t c = 0; c < B->tclass.size(); ++c)
            fprintf(stderr, "pdm_batch_init: tile class %zu: %lld tiles, halo_max=%d, waves %d families %d\n", c,
                    (long long)B->tclass[c].n, B->tclass[c].halo_max, B->tclass[c].waves, B->tclass[c].fams);
    }
    B->inited = true;
    return PDM_OK;
}

int pdm_batch_num_specimens(const pdm_batch* b) { return (b && b->h) ? b->h->nspec : -1; }

int64_t pdm_batch_num_nodes(const pdm_batch* b, int spec)
{
    if (!b || !b->h) return -1;
    if (spec < 0) return b->h->n;
    return (spec < b->h->nspec) ? b->h->members[spec]->n : -1;
}

int64_t pdm_batch_num_entries(const pdm_batch* b, int spec)
{
    if (!b || !b->h) return -1;
    if (spec < 0) return b->h->E;
    return (spec < b->h->nspec) ? b->h->members[spec]->E : -1;
}

int64_t pdm_batch_num_live_entries(const pdm_batch* b, int spec)
{
    if (!b || !b->h) return -1;
    if (spec < 0) return b->h->Elive;
    return (spec < b->h->nspec) ? b->h->members[spec]->Elive : -1;
}

int pdm_batch_prepare_steps(pdm_batch* b, int steps)
{
    if (!b || !b->h) return PDM_E_ARG;
    return pdm_prepare_steps(b->h, steps);
}

int pdm_batch_step_rk4(pdm_batch* b, int nsteps, int* done)
{
    int rc = b ? (b->h && b->h->inited ? PDM_OK : PDM_E_STATE) : PDM_E_ARG;
    if (rc) return rc;
    pdm_handle* B = b->h;
    pdm_handle* const h = B;  // for HCHK
    if (nsteps < 0) return fail(B, PDM_E_ARG, "negative step count");
    HCHK(hipSetDevice(B->dev));
    const int S = B->nspec;
    int total = 0;
    std::vector<std::vector<StepParam>> per(S);
    std::vector<StepParam> flat;
    while (total < nsteps) {
        const int want = std::min(B->opt.graph_chunk, nsteps - total);
        int cnt = 0;
        for (int k = 0; k < S; ++k) {
            plan_steps(B->members[k], want, per[k]);
            cnt = std::max(cnt, (int)per[k].size());
        }
        if (cnt == 0) break;
        // a specimen whose run has ended takes dt = 0 steps, which leave its state unchanged
        flat.assign((size_t)cnt * S, StepParam{0.0, 0.0});
        for (int k = 0; k < S; ++k)
            for (int r = 0; r < cnt; ++r)
                flat[(size_t)r * S + k] = (r < (int)per[k].size()) ? per[k][r] : StepParam{B->members[k]->t, 0.0};
        rc = run_chunk(B, flat);
        if (rc) return rc;
        if (B->opt.energies) {
            // each specimen's energies of its last real step (a dt = 0 step recomputes
            // the elastic energy from four identical trial fields)
            std::vector<double> en(2 * (size_t)cnt * S);
            HCHK(hipMemcpy(en.data(), B->energy, sizeof(double) * en.size(), hipMemcpyDeviceToHost));
            for (int k = 0; k < S; ++k) {
                if (per[k].empty()) continue;
                const size_t r = per[k].size() - 1;
                B->members[k]->ke = en[2 * (r * S + k)];
                B->members[k]->ee = en[2 * (r * S + k) + 1];
            }
        }
        total += cnt;
        if (cnt < want) break;
    }
    if (done) *done = total;
    return PDM_OK;
}

int pdm_batch_get_state(const pdm_batch* b, int spec, double* U, double* V)
{
    int rc = batch_spec(b, spec);
    if (rc) return rc;
    const pdm_handle* B = b->h;
    const pdm_handle* const h = B;  // for HCHK
    const pdm_handle* m = B->members[spec];
    const int64_t off = B->spec_node_ptr[spec], n = m->n;
    std::vector<d4> buf(n);
    HCHK(hipStreamSynchronize(B->stream));
    if (U) {
        HCHK(hipMemcpy(buf.data(), B->U + off, sizeof(d4) * n, hipMemcpyDeviceToHost));
        from_internal(m, buf, U);
    }
    if (V) {
        HCHK(hipMemcpy(buf.data(), B->V + off, sizeof(d4) * n, hipMemcpyDeviceToHost));
        from_internal(m, buf, V);
    }
    return PDM_OK;
}

int pdm_batch_set_state(pdm_batch* b, int spec, const double* U, const double* V)
{
    int rc = batch_spec(b, spec);
    if (rc) return rc;
    pdm_handle* B = b->h;
    pdm_handle* const h = B;  // for HCHK
    const pdm_handle* m = B->members[spec];
    const int64_t off = B->spec_node_ptr[spec], n = m->n;
    std::vector<d4> buf;
    HCHK(hipStreamSynchronize(B->stream));
    if (U) {
        to_internal(m, U, buf);
        HCHK(hipMemcpy(B->U + off, buf.data(), sizeof(d4) * n, hipMemcpyHostToDevice));
        HCHK(hipMemcpy(B->u[0] + off, buf.data(), sizeof(d4) * n, hipMemcpyHostToDevice));
    }
    if (V) {
        to_internal(m, V, buf);
        HCHK(hipMemcpy(B->V + off, buf.data(), sizeof(d4) * n, hipMemcpyHostToDevice));
    }
    return PDM_OK;
}

int pdm_batch_get_clock(const pdm_batch* b, int spec, double* t, int* step, double* dt_next)
{
    int rc = batch_spec(b, spec);
    if (rc) return rc;
    return pdm_get_clock(b->h->members[spec], t, step, dt_next);
}

int pdm_batch_get_energy(const pdm_batch* b, int spec, double* ke, double* ee)
{
    int rc = batch_spec(b, spec);
    if (rc) return rc;
    return pdm_get_energy(b->h->members[spec], ke, ee);
}

int pdm_batch_get_tuning(const pdm_batch* b, int* tile_edge, int* waves_per_tile, int* families_per_wave,
                         double* stage_ms)
{
    if (!b || !b->h) return PDM_E_ARG;
    return pdm_get_tuning(b->h, tile_edge, waves_per_tile, families_per_wave, stage_ms);
}

int pdm_batch_get_spec_tuning(const pdm_batch* b, int spec, int* tile_edge, int* waves_per_tile,
                              int* families_per_wave)
{
    if (!b || !b->h) return PDM_E_ARG;
    const pdm_handle* B = b->h;
    if (spec < 0 || spec >= B->nspec) return PDM_E_ARG;
    int w = B->waves, f = B->fam_per_wave;
    if (!B->tclass.empty()) {
        const TileClass& c = B->tclass[B->spec_class[spec]];
        w = c.waves;
        f = c.fams;
    }
    if (tile_edge) *tile_edge = B->tile_edge[0];
    if (waves_per_tile) *waves_per_tile = w;
    if (families_per_wave) *families_per_wave = f;
    return PDM_OK;
}

int pdm_batch_last_kernel_timing(const pdm_batch* b, double* avg_ms, int64_t* launches)
{
    if (!b || !b->h) return PDM_E_ARG;
    return pdm_last_kernel_timing(b->h, avg_ms, launches);
}

}  // extern "C"
