/*
 * hcc.h -- C ABI of the MI355X (gfx950) implementation of HyperCloudCore.jl's explicit time-step hot path.
 *
 * The reference's user-facing seam for this path is three calls per run (README.md:108-150,
 * test/time_integration.jl:106-111): `ParticleGrid(...)` (src/Grid/ParticleGrids.jl:221-279),
 * `get_time_step(pg, eq, interp)` (src/Grid/ParticleGrids.jl:386-415) and the stepper functor
 * `ts(eq, pg, t, dt)` (src/TimeStepping/TimeSteppers.jl:531-577). Kernel bodies there are Julia closures over
 * user physics, so the replacement plugs in at the operator level: a grid object, a stepper object and a
 * PDE descriptor that selects a compiled kernel variant. Each entry point below names the reference
 * function it replaces.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative code on error; hcc_last_error() gives the text.
 *     No C++ exception crosses this boundary.
 *   - Host pointers are borrowed for the duration of the call only. Device memory is owned by the library.
 *   - States cross the boundary as AoS: M contiguous doubles per node (Julia's Vector{SVector{M,Float64}}).
 *     Positions likewise D doubles per node. Node and slot indices are 0-based here (a Julia binding adds 1).
 *   - One HIP stream per context; calls on one context are not re-entrant (callers serialise per context).
 *   - Only one GPU (the context's device) is used.
 */
#ifndef HCC_H
#define HCC_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hcc_ctx hcc_ctx;
typedef struct hcc_grid hcc_grid;
typedef struct hcc_stepper hcc_stepper;

enum {
    HCC_OK = 0,
    HCC_E_ARG = -1,       /* invalid argument */
    HCC_E_HIP = -2,       /* HIP runtime error */
    HCC_E_UNSUPPORTED = -3,/* PDE / dimension / scheme not compiled in (no fallback path exists) */
    HCC_E_CAPACITY = -4   /* caller buffer too small; required size returned through the size argument */
};

/* ---- context ------------------------------------------------------------------------------------------ */
/* Selects `device`, creates the context's stream. */
int hcc_ctx_create(int device, hcc_ctx** out);
int hcc_ctx_destroy(hcc_ctx* ctx);
/* Text of the last error on this context (or of a failed hcc_ctx_create when ctx is NULL). */
const char* hcc_last_error(const hcc_ctx* ctx);
int hcc_ctx_synchronize(hcc_ctx* ctx);

/* ---- point cloud (host set-up) ------------------------------------------------------------------------ */
/* A rectangular domain as get_rectangular_domain (src/Grid/Domains.jl:46-97) with the ghost canvas of
 * ComputationalDomain (Domains.jl:159-191) for cutoff R = weight-function range. */
typedef struct {
    int32_t D;              /* 1, 2 or 3 */
    double mins[3], maxs[3];
    int32_t periodic[3];
    double dx[3];           /* nominal spacing per axis */
    double R;               /* support radius (ExponentialWeightFunction range) */
    double alpha;           /* ExponentialWeightFunction alpha */
    double jitter[3];       /* `randomness`: uniform jitter amplitude per axis */
    uint64_t seed;          /* seed of the jitter stream (xoshiro256**, splitmix64-seeded) */
} hcc_domain_desc;

/* Replaces get_points + _narrow_band + tags (Domains.jl:220-347) and the Morton reorder
 * (src/Grid/Reordering.jl:55-101). Host code. Writes N nodes (pos: N*D doubles) if cap >= N, else returns
 * HCC_E_CAPACITY. *n_out = N always; canvas (optional, 2*D doubles): canvas mins then maxs;
 * morton_ties (optional): number of equal Morton keys (their order is by original index). */
int hcc_cloud_generate(const hcc_domain_desc* desc, int64_t cap, double* pos, uint8_t* is_boundary, int32_t* tags,
                       int64_t* n_out, double* canvas, int64_t* morton_ties);

/* Seeded uniform doubles in [0,1) from the same generator (initial conditions of the synthetic workloads). */
int hcc_rng_uniform(uint64_t seed, int64_t n, double* out);

/* ---- grid + neighbour search (device) ----------------------------------------------------------------- */
/* Replaces ParticleGrid's device part: uploads the nodes and runs the cell-list neighbour search on the
 * GPU -- periodic wrap, counting-sort bins with descending in-bin order, count, scan, write
 * (src/Grid/NeighborLogic.jl:22-167, 235-329). Positions are wrapped on periodic axes. */
int hcc_grid_create(hcc_ctx* ctx, const hcc_domain_desc* desc, int64_t N, const double* pos, const uint8_t* is_boundary,
                    const int32_t* tags, hcc_grid** out);
int hcc_grid_destroy(hcc_grid* g);
/* total neighbour slots S and the largest count max_nb (pg.meta.max_nb) */
int hcc_grid_num_slots(const hcc_grid* g, int64_t* S, int64_t* max_nb);
/* Row-major neighbour lists as the reference's host layout: ranges N+1 offsets, indices S entries, weights S,
 * distances S*D (any output may be NULL). Same lists in the same order as NeighborLogic.jl:309-329. */
int hcc_grid_get_neighbors(hcc_grid* g, int64_t* ranges, int32_t* indices, double* weights, double* distances);
/* Cell list: bin_of N, cell_start nbins+2, cell_particles N, dims[4] = (dims_x, dims_y, dims_z, nbins). */
int hcc_grid_get_bins(hcc_grid* g, int32_t* bin_of, int64_t* cell_start, int32_t* cell_particles, int64_t* dims);
/* Positions after the periodic wrap (N*D AoS). */
int hcc_grid_get_positions(hcc_grid* g, double* pos);

/* ---- stepper ------------------------------------------------------------------------------------------ */
enum { HCC_PDE_LINADV = 1, HCC_PDE_EULER = 2 };
enum { HCC_LIM_NONE = 0, HCC_LIM_BJ = 1, HCC_LIM_VK = 2 };
enum { HCC_MODE_SOFT = 0, HCC_MODE_HARD = 1 };
/* MOOD strategies (src/TimeStepping/MOOD.jl:7-26): interface orders (o_i, o_j) for EPD0 and StrictEPD0, the edge
 * minimum for EPD1, the minimum over N(i) u N(j) for EPD2; StrictEPD0 also lowers the neighbours of a dropped
 * particle to its order (one hop, no cascade). Criteria (MOOD.jl:140-205): MOODu1 (DMP), MOODPAD (admissibility
 * only), OnlyMOOD (every check flags), MOODu2 (MOODu1 with the plateau test; at MUSCL order 2 the curvature
 * override does not apply). */
enum { HCC_STRAT_NONE = 0, HCC_STRAT_EPD0 = 1, HCC_STRAT_EPD1 = 2, HCC_STRAT_EPD2 = 3, HCC_STRAT_STRICT_EPD0 = 4 };
enum { HCC_CRIT_NONE = 0, HCC_CRIT_U1 = 1, HCC_CRIT_PAD = 2, HCC_CRIT_ONLY = 3, HCC_CRIT_U2 = 4 };
enum { HCC_BC_NONE = 0, HCC_BC_OUTFLOW = 1, HCC_BC_DIRICHLET = 2 };
/* numerical flux (src/Interpolation/FluxFunctions.jl:66-111): RusanovFlux, AlignedRusanovFlux */
enum { HCC_FLUX_RUSANOV = 0, HCC_FLUX_ALIGNED_RUSANOV = 1 };
/* source terms (src/Types.jl:148-172): the built-in decay S_i(U) = -rate_i U (the reference tests' DecaySource,
 * test/unit_schemes.jl:279-284), added to K_E (explicit) or solved in the diagonally implicit stages (implicit) */
enum { HCC_SRC_NONE = 0, HCC_SRC_DECAY_EXPLICIT = 1, HCC_SRC_DECAY_IMPLICIT = 2 };
/* grid movers (src/TimeStepping/GridMovers.jl:14-63): NoGridMover, LagrangianMover (w = fluid velocity),
 * ScaledLagrangianMover(alpha) (w = alpha x fluid velocity). With a mover the particle positions are integrated as
 * stage variables of the same tableau, the flux is taken in each particle's frame (F - w U, wave speeds |lambda - w|),
 * each stage's geometry keeps the step's neighbour lists (weight 0 beyond R), and every step ends with a device
 * neighbour search (periodic wrap) and ShepardVolumes. Not with outflow boundary conditions; linear advection with
 * M = 1 only. The grid's lists and positions belong to the moving stepper: one moving stepper per grid. */
enum { HCC_MOVER_NONE = 0, HCC_MOVER_LAGRANGIAN = 1, HCC_MOVER_SCALED_LAGRANGIAN = 2 };

typedef struct {
    int32_t M;                 /* state components */
    int32_t pde;               /* HCC_PDE_* */
    double gamma;              /* Euler: ratio of specific heats */
    double adv_speed[3][8];    /* linear advection: speed of component m along axis d (diagonal system) */
    int32_t muscl_order;       /* MUSCL max_order: 2, or 3 (advection (1,1), (2,1), (2,2), (3,1), 2-D and 3-D
                                  Euler; buffered RusanovFlux on static grids) */
    int32_t limiter, limiter_mode;
    int32_t mood_strategy, mood_criterion;
    double mood_delta;         /* MOODu1 / MOODu2 d */
    int32_t stages;            /* explicit tableau, s <= 5 */
    double a_t[5][5], b_t[5], c_t[5];
    double substep_ratio;      /* Δt/C per Δt of the MOOD Euler substeps (TimeSteppers.jl:218); 0: check stage candidates */
    int32_t n_bc;
    int32_t bc_tag[16], bc_kind[16];
    double mood_plateau;       /* MOODu2 plateau (fraction of the global range of the previous stage; 0: off) */
    int32_t muscl_streamed;    /* MUSCL(...; streamed = true): divergence accumulated per slot (MUSCL.jl:56, 407-408) */
    int32_t flux;              /* HCC_FLUX_* */
    /* implicit part of an IMEX tableau (IMEXButcherTableau a, b, c, ButcherTableaus.jl:23-45): lower triangular
     * a_im; all zero for an explicit tableau. substep_ratio must be 0 for a tableau with an implicit part. */
    double a_im[5][5], b_im[5], c_im[5];
    int32_t mover;             /* HCC_MOVER_* (pg.mover of the ParticleGrid) */
    double mover_alpha;        /* ScaledLagrangianMover's alpha */
    double mood_ratio;         /* MOODu2's ratio of the nearest curvature magnitudes (MOOD.jl:133, default 0.5; <= 0
                                  off), used by its smooth-extremum override at MUSCL order 3 */
} hcc_stepper_desc;

/* Replaces GeneralRKTimeStepper(pde, MUSCL(...), MOOD(...), tableau) (TimeSteppers.jl:195-229) and the MUSCL
 * geometry caches of prepare_geometry! (src/Interpolation/MUSCL.jl:172-223): inverse moment matrices of the
 * reconstruction and divergence fits, per-slot derivative weights. */
int hcc_stepper_create(hcc_ctx* ctx, hcc_grid* g, const hcc_stepper_desc* desc, hcc_stepper** out);
int hcc_stepper_destroy(hcc_stepper* st);
/* pg.rhos up/download (N*M AoS doubles) */
int hcc_state_upload(hcc_stepper* st, const double* U);
int hcc_state_download(hcc_stepper* st, double* U);
/* Replaces get_time_step(pg, eq, interp) with cached weights (ParticleGrids.jl:386-450): the CFL-1 step. */
int hcc_time_step(hcc_stepper* st, double* dt);
/* Replaces ts(eq, pg, t, dt) (TimeSteppers.jl:531-577): one time step in place on the device state. */
int hcc_step(hcc_stepper* st, double t, double dt);
/* The driver loop's step with its time step taken on the device (test/time_integration.jl:107-111):
 * dt = min(cfl * get_time_step, dt_cap) with Julia's min, then ts(eq, pg, t, dt), queued without a host round trip
 * (pass dt_cap = t_end - t, or +inf). No counterpart in the reference: it removes the host wait of get_time_step from
 * device-resident driver loops. hcc_last_dt returns the dt used (waiting for the step); a CFL step that is not
 * positive and finite is reported by the next call that waits, and the state is then invalid. */
int hcc_step_cfl(hcc_stepper* st, double cfl, double dt_cap);
int hcc_last_dt(hcc_stepper* st, double* dt);
/* Diagnostics of the last step: particle_orders (N), mood_flags (TimeSteppers.jl:258, N bytes),
 * mood_inadmissible (:272), and the number of MOOD recompute iterations. */
int hcc_get_orders(hcc_stepper* st, int32_t* orders);
int hcc_get_mood_flags(hcc_stepper* st, uint8_t* flags);
int hcc_get_inadmissible(hcc_stepper* st, int64_t* count);
int hcc_get_mood_iterations(hcc_stepper* st, int64_t* count);
/* Geometry caches (row-major B*B per node; ok bytes; dweights S*D in row-major slot order, formed by this call
 * from the current geometry: the step's passes keep no copy). B: the reconstruction's basis_count(D, order - 1)
 * terms; the divergence's basis_count(D, 1 or 3) + 1 (MUSCL order 2 or 3), of which only the D derivative rows are
 * kept (the other rows read NaN; a stencil that carries a lower degree has its inverse zero-padded). */
int hcc_get_moments(hcc_stepper* st, double* inv_grad, uint8_t* ok_grad, double* inv_div, uint8_t* ok_div,
                    double* dweights);
/* One full right-hand side (gather + flux pass, orders reset to the maximum) on state Y (N*M AoS):
 * K (N*M) and the limited gradients (N*B*M with B = basis_count(D, order - 1): node-major, then basis term in the
 * order of _generate_exponents, then component). */
int hcc_rhs(hcc_stepper* st, const double* Y, double* K, double* gradients);
/* Timing of the last hcc_step on the context stream (HIP events), milliseconds: the whole step, and the summed
 * durations of its full-domain gather (gradient + limiter) and flux passes, one of each per stage.
 * hcc_step returns once its work is enqueued; this call, the diagnostics above and hcc_step_totals wait for it. */
int hcc_last_step_timing(hcc_stepper* st, double* ms_total, double* ms_gather, double* ms_flux);
/* Sums over the steps since the last reset (reset != 0 clears them after reading): steps, MOOD fixed-point
 * iterations, and the HIP-event milliseconds of hcc_last_step_timing. No counterpart in the reference: it lets a
 * driver time a run of steps without a host round trip per step (benchmark/step.jl times its loop as a whole). */
int hcc_step_totals(hcc_stepper* st, int reset, int64_t* steps, int64_t* mood_iterations, double* ms_total,
                    double* ms_gather, double* ms_flux);
/* No counterpart in the reference. Pass timing: the HIP events around each stage's full gather and flux pass (the
 * gather and flux milliseconds above) are recorded every `every`-th step queued from now on (1, the default: every
 * step; 0: never). Each event between two kernels is a marker packet that leaves the device idle for a few
 * microseconds (the nine of a 2-D Euler RK3 step cost 3.7 % of it at 1 M nodes), so a driver that times a long run
 * may sample. A step without pass events adds 0 to the gather and flux sums. */
int hcc_stepper_set_pass_timing(hcc_stepper* st, int32_t every);
/* The steps of the hcc_step_totals sums whose passes were timed (cleared with them). */
int hcc_step_timed_passes(hcc_stepper* st, int64_t* steps);

/* No counterpart in the reference (its step is a host loop over kernel launches). With on != 0, hcc_step_cfl runs
 * each step as one hipGraph launch: the whole step -- the CFL estimate fused into the stage-1 gather, the gather and
 * flux passes, the stage predictors and each MOOD fixed point as a fixed number of device-gated iterations (an
 * iteration after convergence does nothing) -- captured once per state buffer (two, swapped every step), per (cfl,
 * dt_cap) and per iteration bounds, then replayed with no host decision inside the step. The bounds are the
 * fixed-point iterations the uncaptured steps needed plus one; a captured fixed point that has not converged within its
 * bound sets a device flag, read before the next step is queued or any result is read: that step is then undone and
 * re-run uncaptured (same results), and the bounds grow. Steps that wait for the host by design stay uncaptured: the
 * first step after an upload, hcc_step (host dt), the unfused MOOD strategies and criteria, IMEX stages, moving grids
 * and training-data collection. Results are those of the uncaptured step, bit for bit. hcc_last_step_timing reports
 * 0 for the gather and flux times of a captured step (no events inside the graph). */
int hcc_stepper_set_graph(hcc_stepper* st, int32_t on);
/* Graph launches, captures and re-runs since creation, and the iterations captured per checkpoint (stages values,
 * checkpoints 2 .. stages + 1; may be NULL). */
int hcc_stepper_graph_stats(hcc_stepper* st, int64_t* launches, int64_t* captures, int64_t* reruns, int32_t* bounds);

/* No counterpart in the reference (its KernelAbstractions kernels use one workgroup size, src/Backend.jl:56). Launch
 * choices of the step, none of which changes a result bit: knobs[0] the full passes stage each workgroup's
 * neighbourhood in LDS (1) or read neighbours through the caches (0); knobs[1] lanes per particle in the MOOD fixed
 * point's kernels (kSplit = 8 for D*M <= 8, else 16; or twice it; 0 = kSplit); knobs[2] particles one sweep of the
 * fixed point's grid covers (16384 or 65536); knobs[3] the CFL estimate of hcc_step_cfl fused into the stage-1 gather
 * (1) or its own pass (0); knobs[4] when the first iteration of a fused MOOD fixed point runs in its dense form (plain
 * marks, then the staged flux pass of the marked particles instead of the list kernels): when the same checkpoint of
 * the last step started from at least N / knobs[4] changed particles (knobs[4] > 0), never (0), always (-1). Defaults:
 * staged where the LDS tables exist, kSplit, 65536, fused, 64. */
int hcc_stepper_set_tuning(hcc_stepper* st, const int32_t* knobs);
int hcc_stepper_get_tuning(hcc_stepper* st, int32_t* knobs);
/* The autotuner: times `steps` CFL steps (CFL number cfl) of every launch choice from the stepper's current state,
 * restored before each trial and afterwards, and keeps the fastest if it beats the defaults by more than 2 %. The choice
 * is stored per workload shape (D, M, PDE, flux variant, LDS tables present, log2 of the node count, mean neighbour
 * count, ensemble) for the process: a stepper of a tuned shape takes it without timing (*timed = 0). Outputs (may be
 * NULL): knobs[5] as hcc_stepper_get_tuning, the best and the default milliseconds per step (0 when not timed). The
 * step statistics (hcc_step_totals and the last step's diagnostics) are left as they were. */
int hcc_stepper_autotune(hcc_stepper* st, int32_t steps, double cfl, int32_t* knobs, double* ms_best, double* ms_default,
                         int32_t* timed);

/* The source terms of GeneralRKTimeStepper(pde, interp, mood, explicit_sources, implicit_sources, tableau)
 * (TimeSteppers.jl:195-247): the built-in decay S_i(U) = -rates[i] U with per-particle rates (N doubles, read once),
 * mode HCC_SRC_*. An implicit source with a purely explicit tableau is applied explicitly with the explicit weights
 * (_explicit_implicit_part). No counterpart for arbitrary Julia source closures: their rates are evaluated by the
 * host. */
int hcc_set_source(hcc_stepper* st, int32_t mode, const double* rates);

/* ---- particle volumes, conservation diagnostics, driver loop ------------------------------------------- */
/* Replaces update_volumes!(pg, ShepardVolumes(sweeps)) (src/Grid/Volumes.jl:183-231), which the ParticleGrid
 * constructor runs with one sweep (src/Grid/ParticleGrids.jl:275; hcc_grid_create does the same). sweeps >= 0
 * recomputes the volumes; volumes (N doubles, may be NULL) receives the current ones. The sum of the renormalisation
 * is a left fold in index order. */
int hcc_grid_volumes(hcc_grid* g, int32_t sweeps, double* volumes);
/* Replaces total_mass(pg) / interior_mass(pg) (src/Grid/Volumes.jl:103-117): sum_i V_i U_i over all particles, or
 * the interior ones, a left fold in index order like the reference's mapreduce with init; out: M doubles. */
int hcc_total_mass(hcc_stepper* st, int32_t interior_only, double* out);
/* Replaces solve_equation(ts, eq, pg, tmax, dt; is_cfl, snapshots, remove_ghosts) (test/time_integration.jl:72-132):
 * the driver loop on the device state (dt = dt * get_time_step when is_cfl, capped at tmax - t, stop at dt <= 1e-12),
 * snapshots at the first step reaching each of the snapshots + 1 evenly spaced times and at the end. Outputs (any may
 * be NULL): t_hist[snapshots + 1] (NaN for a snapshot the loop never reaches), x_hist and u_hist
 * [(snapshots + 1) x Ns x D or M] with Ns = N or the interior count (remove_ghosts), mass_hist
 * [(snapshots + 1) x 2M] (total_mass, interior_mass of each snapshot), steps, elapsed (seconds of the loop). */
int hcc_solve(hcc_stepper* st, double tmax, double dt, int32_t is_cfl, int32_t snapshots, int32_t remove_ghosts,
              double* t_hist, double* x_hist, double* u_hist, double* mass_hist, int64_t* steps, double* elapsed);

/* ---- ensembles ---------------------------------------------------------------------------------------- */
/* No counterpart in the reference (one ParticleGrid per stepper); for callers that advance many point clouds of
 * uneven size (ensemble members, parameter sweeps) in one step. Builds a grid holding copies of n instance grids
 * (each from hcc_grid_create, with the same D, R, alpha, minimum spacing and wrap lengths; the parts stay usable):
 * instance k's nodes occupy [offsets[k], offsets[k] + N_k), padded to whole 256-node blocks with inert boundary nodes
 * (no neighbours, tag 0). Every node keeps its instance's neighbour lists in their order and its volume, so a stepper
 * on the ensemble computes each instance exactly as a stepper on its own grid. hcc_step gives every instance the
 * same dt; hcc_step_cfl takes each instance's own CFL step. offsets: n + 1 values (offsets[n] = nodes including the
 * padding; may be NULL). 1 <= n <= 65535. Not with grid movers. */
int hcc_grid_concat(hcc_ctx* ctx, int32_t n, hcc_grid* const* parts, hcc_grid** out, int64_t* offsets);
/* The CFL step (CFL number 1) of every instance of the stepper's grid: n doubles (n = 1 for a grid from
 * hcc_grid_create). */
int hcc_time_steps(hcc_stepper* st, double* dts);

/* ---- learned troubled-node indicator --------------------------------------------------------------------- */
/* No counterpart in the reference (its MOOD is a posteriori only). A small network predicts, before a stage's flux
 * pass, which particles MOOD will drop to first order, so that they take it at once and fewer fixed-point sweeps
 * follow; MOOD's a posteriori check still runs on every candidate and catches what the network misses. The network:
 * 16 features -> 64 -> 64 -> 1 logit, GELU (tanh form), bf16 weights and activations, fp32 accumulation, forward and
 * backward on the matrix cores. Parameters: HCC_MLP_PARAMS floats, W1[16][64], b1[64], W2[64][64], b2[64], w3[64],
 * b3. Feature rows are 16 bf16 values (the upper 16 bits of the float), labels 0/1 bytes. */
#define HCC_MLP_PARAMS 5313
typedef struct hcc_mlp hcc_mlp;
int hcc_mlp_create(hcc_ctx* ctx, hcc_mlp** out);
int hcc_mlp_destroy(hcc_mlp* m);
int hcc_mlp_set_params(hcc_mlp* m, const float* params);   /* also resets the optimiser state */
int hcc_mlp_get_params(hcc_mlp* m, float* params);
/* logits z[n] of n feature rows */
int hcc_mlp_forward(hcc_mlp* m, const uint16_t* X, int64_t n, float* z);
/* gradient (HCC_MLP_PARAMS floats) and value of mean_s w_s (softplus(z_s) - y_s z_s), w_s = pos_weight where
 * y_s = 1, else 1 */
int hcc_mlp_gradients(hcc_mlp* m, const uint16_t* X, const uint8_t* y, int64_t n, float pos_weight, float* grads,
                      double* loss);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP graph capture,
 * tight training loops, harnesses that forbid hipMalloc): hcc_mlp_gradients on device-resident data, allocating
 * nothing. X (n x 16 bf16 rows) and y (n label bytes) are device pointers; grads (device, HCC_MLP_PARAMS + 1 floats)
 * receives the gradient, then the loss. The work is queued on the context stream and the call returns without waiting.
 * The workgroup partials go to a static device buffer of HCC_MLP_NOALLOC_SCRATCH_ROWS x (HCC_MLP_PARAMS + 1) floats
 * (10.4 MiB, one per GPU, held until the process ends), resolved by hcc_mlp_create; every batch size fits, and a
 * launch that would need more rows is refused with HCC_E_CAPACITY. Calls that overlap on one GPU (different contexts
 * or threads) race on that buffer. Same bits as hcc_mlp_gradients (the same kernels and reduction order). */
#define HCC_MLP_NOALLOC_SCRATCH_ROWS 512
int hcc_mlp_gradients_noalloc(hcc_mlp* m, const uint16_t* X, const uint8_t* y, int64_t n, float pos_weight,
                              float* grads);
/* epochs of Adam (beta 0.9 / 0.999, eps 1e-8) over mini-batches of `batch` rows in the given order; losses (may be
 * NULL): the mean batch loss of each epoch */
int hcc_mlp_train(hcc_mlp* m, const uint16_t* X, const uint8_t* y, int64_t n, int32_t epochs, int32_t batch, float lr,
                  float pos_weight, double* losses);

/* The indicator in the step (2-D Euler): after each stage's full gather, which then also stores 16 features per
 * particle, particles still at the full order whose logit exceeds `threshold` take first order before the flux pass;
 * MOOD's a posteriori check of every candidate still runs. m = NULL switches it off (the step is the plain one). The
 * network stays owned by the caller and must outlive its use. */
int hcc_set_indicator(hcc_stepper* st, const hcc_mlp* m, float threshold);
/* Training data: with on, each step keeps every stage's features and the orders before that stage's flux pass and
 * after its checkpoint (the label of a particle at the full order: MOOD dropped it). */
int hcc_indicator_collect(hcc_stepper* st, int32_t on);
/* The last step's stage (0-based): feats N x 16 bf16, pre and post orders N each (any may be NULL). */
int hcc_get_indicator_samples(hcc_stepper* st, int32_t stage, uint16_t* feats, int8_t* pre, int8_t* post);
/* Particles the indicator lowered in the last step. */
int hcc_get_indicator_count(hcc_stepper* st, int64_t* count);

#ifdef __cplusplus
}
#endif
#endif /* HCC_H */
