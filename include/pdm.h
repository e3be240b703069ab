/*
 * pdm.h -- C ABI of the MI355X (gfx950) linear bond-based peridynamics library.
 *
 * This is the drop-in boundary for the hot path of the reference solver:
 * the pdSolve::InitialProblem / SolveDynamic / SolveStatic drivers and the
 * pdDataManager calls they make.  Each entry point below names the reference
 * interface it replaces (file:line).  Plain pointers and sizes only.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative PDM_E* code on error;
 *    pdm_last_error() then returns a message.  (The reference calls exit(0) or
 *    throws on the same conditions, pdDataManager.h:400-419, pdDataManager.cpp:980-992.)
 *  - 3-vectors are node-major AoS (x,y,z of node 0, node 1, ...), the reference's
 *    _rowOfNode/displStep layout (pdDataManager.cpp:1085); node ids are the
 *    reference's lattice ids (x fastest, pdDataManager.cpp:846-909).
 *  - Host buffers are caller-owned and copied in/out.  All device memory is
 *    owned by the handle.  One HIP stream per handle; calls on one handle must
 *    be serialised by the caller, distinct handles are independent.
 */
#ifndef PDM_H
#define PDM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PDM_OK 0
#define PDM_E_ARG -1        /* bad argument / null pointer */
#define PDM_E_IO -2         /* input file missing or unreadable */
#define PDM_E_PARSE -3      /* required keyword missing, value missing, bad value */
#define PDM_E_MODEL -4      /* node outside every material space, unsupported mode */
#define PDM_E_HIP -5        /* HIP runtime error */
#define PDM_E_STATE -6      /* call out of order (e.g. step before init) */
#define PDM_E_LIMIT -7      /* a size limit of this build exceeded */
#define PDM_E_CONVERGE -8   /* static solve did not reach the tolerance */

/* precision of the bond-force gather and of CG matvecs */
#define PDM_FP64 0          /* reference arithmetic in FP64 */
#define PDM_FP32C 1         /* FP32 bond force (cancellation-free stretch), FP32 family sums,
                               FP64 state; CG matvecs on FP32 weights */
#define PDM_FP32C2 2        /* as PDM_FP32C with two-float (hi, lo) FP32 family sums (TwoSum of
                               FP32 partials of 4 bond forces); CG matvecs on FP32 weights with
                               a 16-bit significand extension (39-bit weights) */

typedef struct pdm_handle pdm_handle;

typedef struct {
    int precision;          /* PDM_FP64 or PDM_FP32C */
    int device;             /* HIP device ordinal */
    int use_graph;          /* capture each chunk of RK4 steps in a hipGraph (default 1) */
    int graph_chunk;        /* steps per captured graph (default 50) */
    int energies;           /* compute KE/EE every step (reference behaviour, default 1) */
    int tile_edge;          /* tile edge in grid spacings, 2..7 */
    int waves_per_tile;     /* wavefronts per tile workgroup, 4 or 8 */
    int families_per_wave;  /* families a wavefront sums at once, 1, 2 or 4; with all three 0 the
                               layout is autotuned at pdm_init (specimens with >= 4M family
                               entries; smaller ones take 4, 8, 2) */
    int verbose;
} pdm_options;

typedef struct {
    int iterations;         /* PCG iterations */
    int refinements;        /* FP64 residual evaluations (each replaces the CG residual) */
    double rel_residual;    /* final ||b - K u|| / ||b|| in FP64 */
    double seconds;         /* wall time of the solve */
} pdm_static_stats;

/* Parsed deck values (pdDataManager getters, pdDataManager.h:104-142). */
typedef struct {
    int grid_dimension[3];
    double grid_spacing, horizon, grid_center[3];
    double total_time;
    int total_step, step_dump_frequency;
    double safety_factor, fixed_time_step;   /* fixed_time_step < 0: stable dt */
    int num_materials, num_bcs, num_material_spaces, num_deletion_spaces, num_boundary_spaces, num_slits;
    double damping_coeff;
    int bdryeff_on, cci_on, cci_modified, fgi_on, ai_on, mls_on, intg_point;
    double support_domain_size, tolerance;
    double jitter;
    long long jitter_seed;
    int traction_body;
} pdm_problem_info;

/* Fill *opt with defaults. */
int pdm_options_default(pdm_options* opt);

/* Parse a pds_input.txt deck (CRLF or LF) and create a handle; no GPU work.
 * Replaces pdDataManager::ReadInputFile (pdDataManager.cpp:4-758) with
 * ReadKeyword/FindKeyword (pdDataManager.h:270-421) and SetWkDir's input
 * check (pdDataManager3.cpp:1503-1541).  Integration: CCI (cci_on 1) in any precision;
 * fixed Gaussian integration (cci_on 0, fgi_on 1, intg_point 1..6 or 8) in PDM_FP64;
 * adaptive integration (cci_on 0, fgi_on 0, ai_on 1 mls, intg_point 1..5) in PDM_FP64, without
 * error control (error_control_on 1 is refused with PDM_E_MODEL). */
int pdm_create(const char* input_path, const pdm_options* opt, pdm_handle** out);

/* Same, from deck text in memory. */
int pdm_create_from_text(const char* deck_text, const pdm_options* opt, pdm_handle** out);

/* Parsed deck values of a created handle (before or after pdm_init). */
int pdm_get_problem_info(const pdm_handle* h, pdm_problem_info* info);

/* Release the handle and its device memory (pdDataManager::~pdDataManager, pdDataManager3.cpp:7-18). */
int pdm_destroy(pdm_handle* h);

/* Message of the last failed call on h (or of the last failed pdm_create if h is NULL). */
const char* pdm_last_error(const pdm_handle* h);

/* Select the device, then the InitialProblem sequence on the GPU (pdSolve.cpp:10-22): InitNode
 * (pdDataManager.cpp:813-1098), InitBond (pdDataManager2.cpp:12-70, cell-binned
 * GPU search, bond set identical to the reference's), InitMaterial
 * (pdDataManager.cpp:1297-1338), InitBdryEffectFactor (pdDataManager.cpp:1340-1473),
 * plus per-entry force coefficients and the initial state. */
int pdm_init(pdm_handle* h);

/* Sizes (pdDataManager::GetNumNode / GetNumBond, pdDataManager.h:107-108). */
int64_t pdm_num_nodes(const pdm_handle* h);
int64_t pdm_num_bonds(const pdm_handle* h);     /* unique bonds */
int64_t pdm_num_entries(const pdm_handle* h);   /* directed family entries = 2 x bonds */
int pdm_max_family(const pdm_handle* h);        /* _numMaxFamily, pdDataManager.cpp:1345-1387 */
/* Family entries the force gathers evaluate.  Under CCI a cut entry and an entry with
 * beta = 0 (r >= delta + dx/2) contribute exactly 0 to the force and the elastic energy
 * (CalcBondForceCCI, pdDataManager2.cpp:1293-1351, scales both by beta), so the gathers
 * skip them; the body force still counts every unbroken entry.  Under fixed Gaussian
 * integration each Gauss point has its own beta: the uncut entries with at least one
 * point of nonzero beta (see pdm_num_live_gauss_pairs).  No reference counterpart (the
 * reference evaluates all pdm_num_entries). */
int64_t pdm_num_live_entries(const pdm_handle* h);

/* Node data: X (3N, may be NULL), material index (N, may be NULL),
 * boundary-condition index (N, -1 = none; may be NULL).  pdNode getters, pdNode.h:27-142. */
int pdm_get_nodes(const pdm_handle* h, double* X, int32_t* mat, int32_t* bc);

/* Family CSR in reference order: row_ptr (N+1), col (E, ascending within a row,
 * the order InitBond builds families in), cut (E: 1 = bond crosses a slit, the
 * pdBond::_break state it starts with).  Any pointer may be NULL. */
int pdm_get_csr(const pdm_handle* h, int64_t* row_ptr, int32_t* col, uint8_t* cut);

/* Boundary-effect factor per node (pdNode::GetBdryEffFactor). */
int pdm_get_bdryeff(const pdm_handle* h, double* be);

/* Stable time step (pdDataManager::CalcStableTimeStep, pdDataManager2.cpp:6-10) and min bond length. */
int pdm_get_timestep(const pdm_handle* h, double* dt, double* min_bond_length);

/* Family force sum F_i = sum_j f_ij for a displacement field u (3N host), every
 * node's displacement taken from u: the CCI family loop of CalcBondBasedMotionRK4
 * (pdDataManager.cpp:1795-1864) with CalcBondForceCCI (pdDataManager2.cpp:1293-1351).
 * Runs the handle's precision.  elast (N, may be NULL) receives the per-node energy density.
 * Fixed Gaussian integration: the sum over each neighbour's Gauss points
 * (pdDataManager.cpp:1918-1957), every Gauss point interpolated from u
 * (CalcIntgPointDisplRK4's non-BC branch, pdDataManager.cpp:2426-2479). */
int pdm_compute_forces(pdm_handle* h, const double* u, double* F, double* elast);

/* pdm_compute_forces for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops): u (3N), F (3N) and elast (N, may be NULL) are DEVICE pointers in the
 * same reference-order layout, and the call makes no device allocation: its internal-order
 * copies live in a per-GPU scratch array of the library sized for 524,288 nodes (36 MiB on each
 * GPU used, held until the process ends; larger problems get PDM_E_LIMIT).  The work is
 * enqueued on the handle's stream; with sync != 0 the call waits for it, else the caller
 * orders it (e.g. captures the stream).  Calls on one GPU share its scratch: calls that
 * overlap on one GPU race.  CCI decks only.  No counterpart in the reference: it serves the
 * callers that cannot allocate. */
int pdm_compute_forces_noalloc(pdm_handle* h, const double* u, double* F, double* elast, int sync);

/* Areal force density at the given nodes (reference node ids): the post-processing of
 * pdSolve::PostAFD / pdDataManager::CalcArealForceDensity (pdDataManager3.cpp:442-620).  For
 * normal e_n, the collinear lattice nodes x_s - (1 + k) dx e_n (k < int(delta / dx), exact
 * coordinate equality) each add ds = delta / int(delta / dx) times the CCI bond forces of
 * their family members in the half space (x_j - x_s) . e_n >= 0; sigma[9 t + 3 n + c].
 * From the current displacement U (pdm_set_state loads a dump's); bdryeff (N, may be NULL)
 * replaces the boundary-effect factors, as PostAFD takes them from the dump.  Bonds cut by a
 * slit (extension) carry no force.  CCI decks. */
int pdm_areal_force_density(pdm_handle* h, const int64_t* nodes, int64_t count, const double* bdryeff,
                            double* sigma);

/* Fixed Gaussian integration: Gauss points per node (n^3; 0 for CCI), and their positions
 * and current displacements, [node][point][xyz] in reference node order with points in
 * SetGaussianPoint's order (pdDataManager2.cpp:72-104; pdNode::GetGaussianPoint and the
 * pdGauss getters, pdNode.h:100-102, pdGauss.h).  Xg or Ug may be NULL. */
int pdm_num_gauss_points(const pdm_handle* h);
int pdm_get_gauss(const pdm_handle* h, double* Xg, double* Ug);
/* Fixed Gaussian integration: (entry, Gauss point) pairs the gathers evaluate, those of
 * uncut entries whose point has beta > 0 (0 for CCI).  A pair with beta = 0 adds exactly 0
 * to the force and the energy (pdDataManager.cpp:1918-1957 scales both by beta).  No
 * reference counterpart (the reference evaluates every pair); for throughput accounting. */
int64_t pdm_num_live_gauss_pairs(const pdm_handle* h);

/* Adaptive integration (ai_on 1): the trapezoidal points of every family entry, node j's
 * cell clipped to node i's horizon sphere (CalcBondForceIntg, pdDataManager2.cpp:1222-1291;
 * InitIntegrationPoint's AI branch, :257-297; pdIntgManager.cpp:49-451, 523-3482), 0 for the
 * other modes.  pdm_get_ai_points: per point its entry's node i and node j (reference ids) and
 * x, y, z relative to node i and the weight (xw, 4 per point); the points of an entry are
 * contiguous, in x (outer), y, z (inner) order; entries are in no particular order.  Any
 * pointer may be NULL. */
int64_t pdm_num_ai_points(const pdm_handle* h);
/* Adaptive integration with MLS: the stored interpolation coefficients (one per point and
 * contributing node of the point's cell owner), 0 otherwise; for throughput accounting. */
int64_t pdm_num_ai_coefficients(const pdm_handle* h);
int pdm_get_ai_points(const pdm_handle* h, int32_t* node, int32_t* nbr, double* xw);

/* Set / get the dynamic state U, V (3N each, host; NULL skips). */
int pdm_set_state(pdm_handle* h, const double* U, const double* V);
int pdm_get_state(const pdm_handle* h, double* U, double* V);

/* Advance nsteps steps of pdSolve::SolveDynamic's loop (pdSolve.cpp:55-120):
 * t += dt, InitDynamicSolve (pdDataManager2.cpp:410-439), RK4
 * (pdDataManager.cpp:1994-2134) with UpdateNodeStatusWithBC each stage
 * (pdDataManager.cpp:2497-2602), shortened final step.  Stops early (returning
 * the count in *done) when total_step or total_time is reached. */
int pdm_step_rk4(pdm_handle* h, int nsteps, int* done);

/* Capture and cache the HIP graph of a `steps`-step chunk without running it, so that a later
 * pdm_step_rk4 whose chunks have that length (options.graph_chunk >= steps) replays it with no
 * capture on its path; two chunk lengths stay cached per handle.  No counterpart in the
 * reference (which launches no graphs); for timed loops and serving loops. */
int pdm_prepare_steps(pdm_handle* h, int steps);

/* Current time, step count, kinetic and elastic energy of the last step
 * (GetKinetTotal / GetElastStep, pdDataManager3.cpp:1415-1428). */
int pdm_get_clock(const pdm_handle* h, double* t, int* step, double* dt_next);
int pdm_get_energy(const pdm_handle* h, double* ke, double* ee);

/* The whole dynamic driver with the reference's files in out_dir:
 * pds_output.dat lines, pds_tecplot_<step>.dat at step 1 / every dump / last
 * step (pdSolve.cpp:24-130, 708-732). */
int pdm_solve_dynamic(pdm_handle* h, const char* out_dir);

/* Static equilibrium K u = b on the tangent of the same bond operator (CCI decks), matrix-free
 * Jacobi-preconditioned CG (matvecs in the handle's precision) whose residual is
 * replaced by the FP64 residual every 25 iterations, to rel_tol in FP64.  Replaces the
 * dense SolveStatic/AssembleEquation/Gauss path (pdSolve.cpp:132-673), which is
 * broken in the reference.  On success U holds the solution and V = 0. */
int pdm_solve_static(pdm_handle* h, double rel_tol, int max_iterations, pdm_static_stats* stats);

/* Apply the tangent operator y = K u (3N host each), in the handle's precision (CCI decks). */
int pdm_apply_tangent(pdm_handle* h, const double* u, double* y);

/* Tecplot dump, same line format as pdSolve::WriteTecplot / pdNode::Print
 * (pdSolve.cpp:708-732, pdNode.cpp:418-433). */
int pdm_write_tecplot(const pdm_handle* h, const char* path, int step);

/* Tile layout of the gather kernels: tile edge (grid spacings), wavefronts per tile,
 * families per wavefront, and, when the autotuner chose them (all three options 0 and
 * at least 4M family entries), the measured time of one stage launch in ms (else 0).
 * No counterpart in the reference; it reports the GPU layout choice. */
int pdm_get_tuning(const pdm_handle* h, int* tile_edge, int* waves_per_tile, int* families_per_wave,
                   double* stage_ms);

/* The RK4 stage kernel the launches take (CCI decks): its register budget per lane (128, or 80
 * for the FP32 operations on one-material specimens when the LDS holds a third tile per CU) and
 * its LDS bytes per tile workgroup.  No counterpart in the reference; it reports the GPU
 * launch choice. */
int pdm_get_stage_launch(const pdm_handle* h, int* max_vgprs, int64_t* lds_bytes);

/* Time of the last kernel launches, from HIP events on the handle's stream:
 * average duration in ms of the dominant (bond-force stage) kernel over the last
 * pdm_step_rk4 call, and the number of launches it averaged. */
int pdm_last_kernel_timing(const pdm_handle* h, double* avg_ms, int64_t* launches);

/* ------------------------------------------------------------------ batches
 * Many specimens (decks) stepped together: every RK4 stage is one launch over all
 * specimens' tiles, scheduled by a CSR of specimens over the tile list; each specimen
 * keeps its own deck, time step, clock, boundary conditions and energies.  A specimen
 * whose run (total_step / total_time) has ended takes steps of dt = 0, which leave its
 * state unchanged.  The reference has no batch driver; a batch replaces running
 * pdSolve::InitialProblem + SolveDynamic (pdSolve.cpp:10-130) once per deck. */
typedef struct pdm_batch pdm_batch;

/* Parse nspec decks (files, or texts in memory); no GPU work.  Options are shared.  CCI decks only. */
int pdm_batch_create(const char* const* input_paths, int nspec, const pdm_options* opt, pdm_batch** out);
int pdm_batch_create_from_text(const char* const* deck_texts, int nspec, const pdm_options* opt, pdm_batch** out);
int pdm_batch_destroy(pdm_batch* b);
const char* pdm_batch_last_error(const pdm_batch* b);

/* InitialProblem for every specimen, then one combined tile layout (autotuned when the
 * layout options are 0 and the batch holds at least 4M family entries). */
int pdm_batch_init(pdm_batch* b);

int pdm_batch_num_specimens(const pdm_batch* b);
int64_t pdm_batch_num_nodes(const pdm_batch* b, int spec);    /* spec < 0: all specimens */
int64_t pdm_batch_num_entries(const pdm_batch* b, int spec);  /* spec < 0: all specimens */
int64_t pdm_batch_num_live_entries(const pdm_batch* b, int spec);  /* as pdm_num_live_entries */

/* Advance every specimen by up to nsteps steps of its own SolveDynamic loop; *done =
 * the largest count any specimen took. */
int pdm_batch_step_rk4(pdm_batch* b, int nsteps, int* done);
/* pdm_prepare_steps for the batch's launches (no counterpart in the reference). */
int pdm_batch_prepare_steps(pdm_batch* b, int steps);

/* Per-specimen state (3N, reference node order), clock and energies, as the single-
 * specimen calls. */
int pdm_batch_get_state(const pdm_batch* b, int spec, double* U, double* V);
int pdm_batch_set_state(pdm_batch* b, int spec, const double* U, const double* V);
int pdm_batch_get_clock(const pdm_batch* b, int spec, double* t, int* step, double* dt_next);
int pdm_batch_get_energy(const pdm_batch* b, int spec, double* ke, double* ee);
/* The batch's layout: the tile edge, and the waves and families of the tile class with the
 * most tiles (tiles are launched in classes by how many fit a CU's LDS, each with its own
 * layout); stage_ms: the autotuned stage time summed over the classes. */
int pdm_batch_get_tuning(const pdm_batch* b, int* tile_edge, int* waves_per_tile, int* families_per_wave,
                         double* stage_ms);
/* The layout specimen spec's tiles run with (its class's).  Batches only: no counterpart in
 * the reference. */
int pdm_batch_get_spec_tuning(const pdm_batch* b, int spec, int* tile_edge, int* waves_per_tile,
                              int* families_per_wave);
int pdm_batch_last_kernel_timing(const pdm_batch* b, double* avg_ms, int64_t* launches);

#ifdef __cplusplus
}
#endif
#endif
