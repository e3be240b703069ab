// pdm_api.cpp -- the C ABI (include/pdm.h): handle lifecycle, InitialProblem on the GPU,
// the RK4 driver with hipGraph-captured steps, state transfer and output writers.
#include <algorithm>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>

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
    if (!d.cci_on) {
        g_create_err = d.fgi_on ? "fixed Gaussian integration (fgi_on) is not supported by this build"
                                : "adaptive integration (ai_on) is not supported by this build";
        delete h;
        return PDM_E_MODEL;
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
    return PDM_OK;
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
    }
    if (V) {
        to_internal(h, V, buf);
        HCHK(hipMemcpy(h->V, buf.data(), sizeof(d4) * n, hipMemcpyHostToDevice));
    }
    return PDM_OK;
}

int alloc_parts(pdm_handle* h);

int ensure_chunk(pdm_handle* h, int cap)
{
    if (h->chunk_cap >= cap) return PDM_OK;
    dfree(h->steps);
    dfree(h->energy);
    HCHK(hipMalloc(&h->steps, sizeof(StepParam) * cap));
    HCHK(hipMalloc(&h->energy, sizeof(double) * 2 * cap));
    h->chunk_cap = cap;
    if (h->graph_exec) {
        (void)hipGraphExecDestroy(h->graph_exec);
        h->graph_exec = nullptr;
        h->graph_steps = 0;
    }
    return h->tiles ? alloc_parts(h) : PDM_OK;  // one row of tile partials per step
}

// Enqueue `count` steps whose parameters sit in h->steps[0..count).
int enqueue_steps(pdm_handle* h, int count)
{
    StageArgs a = stage_args(h);
    for (int k = 0; k < count; ++k) {
        for (int st = 0; st < 4; ++st) {
            a.uin = h->u[st & 1];
            a.uout = h->u[(st + 1) & 1];
            HCHK(launch_stage(st, h->opt.precision, a, k, h->stream));
        }
    }
    // per-step KE / EE totals, one launch for the chunk
    if (a.energies)
        HCHK(launch_energy_steps(h->part, h->ntl, h->spec_tile_ptr_d, h->nspec, count, h->energy, h->stream));
    return PDM_OK;
}

int run_chunk(pdm_handle* h, const std::vector<StepParam>& sp)
{
    int count = (int)sp.size();
    if (count == 0) return PDM_OK;
    int rc = ensure_chunk(h, std::max(count, h->opt.graph_chunk));
    if (rc) return rc;
    HCHK(hipMemcpyAsync(h->steps, sp.data(), sizeof(StepParam) * count, hipMemcpyHostToDevice, h->stream));
    bool graph = h->opt.use_graph && count == h->opt.graph_chunk;
    if (graph && !h->graph_exec) {
        hipGraph_t g;
        HCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue_steps(h, count);
        hipError_t ce = hipStreamEndCapture(h->stream, &g);
        if (rc) return rc;
        HCHK(ce);
        HCHK(hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0));
        HCHK(hipGraphDestroy(g));
        h->graph_steps = count;
    }
    HCHK(hipEventRecord(h->ev0, h->stream));
    if (graph) HCHK(hipGraphLaunch(h->graph_exec, h->stream));
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
        double e[2];
        HCHK(hipMemcpy(e, h->energy + 2 * (count - 1), sizeof e, hipMemcpyDeviceToHost));
        h->ke = e[0];
        h->ee = e[1];
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
void free_tiles(pdm_handle* h)
{
    dfree(h->permd); dfree(h->Xi); dfree(h->u[0]); dfree(h->u[1]); dfree(h->rpi); dfree(h->slot);
    dfree(h->self_slot); dfree(h->coef64); dfree(h->coef32); dfree(h->mati); dfree(h->bci); dfree(h->nb);
    dfree(h->tiles); dfree(h->tref); dfree(h->tile_start); dfree(h->halo_ptr); dfree(h->halo);
    dfree(h->part);
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
    HCHK(hipStreamSynchronize(h->stream));
    return PDM_OK;
}

// Autotuner: time whole RK4 steps on this specimen for every candidate tile edge,
// wavefronts per tile and families per wavefront, keep the fastest.  The stage
// kernel's speed depends on the halo-to-tile ratio (tile edge), on the LDS and
// register footprint that sets the tiles resident per CU, and on how many lanes share
// a family; the best point differs between precisions and horizons.
int autotune(pdm_handle* h)
{
    static const int kEdges[] = {4, 5, 6};
    static const int kWaves[] = {8, 4};
    static const int kFams[] = {2, 4};
    double best = 1e300;
    int bt = 0, bw = 8, bf = 2;
    int rc = ensure_chunk(h, 4);
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
        const StepParam sp[4] = {{h->dt_stable, h->dt_stable}, {h->dt_stable, h->dt_stable},
                                 {h->dt_stable, h->dt_stable}, {h->dt_stable, h->dt_stable}};
        HCHK(hipMemcpyAsync(h->steps, sp, sizeof sp, hipMemcpyHostToDevice, h->stream));
        for (int w : kWaves)
            for (int f : kFams) {
                h->waves = w;
                h->fam_per_wave = f;
                rc = enqueue_steps(h, 1);  // warm-up (first launch sets the LDS limit)
                if (rc) return rc;
                HCHK(hipEventRecord(h->ev0, h->stream));
                rc = enqueue_steps(h, 3);
                if (rc) return rc;
                HCHK(hipEventRecord(h->ev1, h->stream));
                HCHK(hipEventSynchronize(h->ev1));
                float ms = 0.f;
                HCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
                const double per = ms / 12.0;
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

}  // namespace

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
    if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
    dfree(h->Xd); dfree(h->rp); dfree(h->col); dfree(h->cut); dfree(h->be); dfree(h->matd);
    dfree(h->permd); dfree(h->Xi); dfree(h->u[0]); dfree(h->u[1]); dfree(h->rpi); dfree(h->slot);
    dfree(h->self_slot); dfree(h->coef64); dfree(h->coef32); dfree(h->mati); dfree(h->bci); dfree(h->nb);
    dfree(h->tiles); dfree(h->tref); dfree(h->tile_start); dfree(h->halo_ptr); dfree(h->halo);
    dfree(h->U); dfree(h->V); dfree(h->Vb); dfree(h->k1); dfree(h->k23);
    dfree(h->elast); dfree(h->part); dfree(h->energy); dfree(h->steps);
    dfree(h->specs_d); dfree(h->mats_d); dfree(h->bcs_d); dfree(h->tile_spec_d); dfree(h->spec_tile_ptr_d);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
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
    if (h->opt.tile_edge) {
        if (h->opt.tile_edge < 2 || h->opt.tile_edge > 7) return fail(h, PDM_E_ARG, "tile_edge must be 2..7");
        h->tile_edge[0] = h->tile_edge[1] = h->tile_edge[2] = h->opt.tile_edge;
    }
    if (h->opt.waves_per_tile) {
        if (h->opt.waves_per_tile != 4 && h->opt.waves_per_tile != 8) return fail(h, PDM_E_ARG, "waves_per_tile must be 4 or 8");
        h->waves = h->opt.waves_per_tile;
    }
    if (h->opt.families_per_wave) {
        int f = h->opt.families_per_wave;
        if (f != 1 && f != 2 && f != 4) return fail(h, PDM_E_ARG, "families_per_wave must be 1, 2 or 4");
        h->fam_per_wave = f;
    }
    // CalcStableTimeStep (pdDataManager2.cpp:6-10)
    h->dt_stable = d.fixed_dt_on ? d.fixed_dt : d.safety * h->min_bond / h->max_c;
    int64_t n = std::max<int64_t>(h->n, 1);
    for (d4** p : {&h->U, &h->V, &h->Vb, &h->k1, &h->k23}) {
        HCHK(hipMalloc(p, sizeof(d4) * n));
        HCHK(hipMemsetAsync(*p, 0, sizeof(d4) * n, h->stream));
    }
    HCHK(hipMalloc(&h->elast, sizeof(double) * n));
    HCHK(hipMemsetAsync(h->elast, 0, sizeof(double) * n, h->stream));
    const bool tune = !h->opt.tile_edge && !h->opt.waves_per_tile && !h->opt.families_per_wave &&
                      h->E >= kTuneMinEntries;
    rc = tune ? autotune(h) : build_tiles(h);
    if (rc) return rc;
    rc = upload_spec(h);
    if (rc) return rc;
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
                stage_lds_bytes(tile_view(h), !h->uniform_ecrit, true, h->opt.precision != PDM_FP64));
    // the driver's clock
    h->dt_cur = h->dt_stable;
    h->t = 0.0;
    h->step = 0;
    h->last_step = d.total_step;
    h->inited = true;
    return PDM_OK;
}

int64_t pdm_num_nodes(const pdm_handle* h) { return h ? h->n : -1; }
int64_t pdm_num_entries(const pdm_handle* h) { return h ? h->E : -1; }
int64_t pdm_num_bonds(const pdm_handle* h) { return h ? h->E / 2 : -1; }
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
    HCHK(launch_forces(h->opt.precision, tile_view(h), ud, Fd, ed, h->stream));
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

int pdm_apply_tangent(pdm_handle* h, const double* u, double* y)
{
    int rc = require_init(h);
    if (rc) return rc;
    if (!u || !y) return fail(h, PDM_E_ARG, "null u or y");
    int64_t n = h->n;
    if (n == 0) return PDM_OK;
    std::vector<d4> buf;
    to_internal(h, u, buf);
    d4 *ud = nullptr, *yd = nullptr;
    HCHK(hipMalloc(&ud, sizeof(d4) * n));
    HCHK(hipMalloc(&yd, sizeof(d4) * n));
    HCHK(hipMemcpy(ud, buf.data(), sizeof(d4) * n, hipMemcpyHostToDevice));
    HCHK(launch_tangent(h->opt.precision, tile_view(h), ud, yd, h->stream));
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
    std::vector<StepParam> sp;
    std::vector<double> en;
    // break chunks at steps that write a dump (1, every dump, last)
    for (;;) {
        int next = h->step + 1;
        int want = 1;
        if (next != 1) {
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
    HCHK(hipSetDevice(h->dev));
    return solve_static(h, rel_tol, max_iterations, stats);
}

}  // extern "C"
