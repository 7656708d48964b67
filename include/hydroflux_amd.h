/*
 * hydroflux_amd: C ABI of the MI355X (gfx950) solver-2d second-order step.
 *
 * Drop-in boundary for hydroflux's solver-2d time step. Each entry point
 * replaces one item of the reference's public Rust API (solver-2d/src/lib.rs:
 * 74-91); the comment on each names the reference item (file:line) whose
 * behaviour it reproduces. Grids stay resident in device memory between
 * calls; host arrays cross the boundary only at upload/download.
 *
 * Layout at the boundary: row-major (n_rows, n_cols), row i along y, column
 * j along x (solver-2d/src/lib.rs:11-25). States are AoS float64 triples
 * (h, hu, hv) per cell, the logical order of the reference's
 * Array2<Conserved2D>; the Dual kind takes a second array of the same shape
 * holding the tangents (dval) of (h, hu, hv). Plain pointers and sizes only.
 *
 * Errors are returned as int status codes instead of the reference's panics
 * and SimError values (sim.rs:81-115); hf2d_last_error() gives the message.
 * One handle per host thread; calls on one handle are serialised by the
 * caller.
 */
#ifndef HYDROFLUX_AMD_H
#define HYDROFLUX_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. 1-4 map 1:1 onto SimError (sim.rs:81-115). */
enum hf2d_status {
  HF2D_OK = 0,
  HF2D_SHAPE_MISMATCH = 1,         /* SimError::ShapeMismatch / shape panics */
  HF2D_INVALID_CONFIG = 2,         /* SimError::InvalidConfig / Mesh2DG::new panics */
  HF2D_DEGENERATE_DT = 3,          /* SimError::DegenerateDt */
  HF2D_STEP_BUDGET_EXHAUSTED = 4,  /* SimError::StepBudgetExhausted */
  HF2D_HIP_ERROR = 5,              /* device runtime failure */
  HF2D_INVALID_ARGUMENT = 6,       /* null handle / pointer, unsupported kind */
  HF2D_CAPACITY = 7                /* batch: no free slot or the scenario does not fit in HBM */
};

/* Scalar of the grid: f64, forward-mode Dual (autograd/src/dual.rs:12), the
 * FP32 precision form (no reference counterpart: FP32 state and flux
 * arithmetic, bed as an FP64 datum per 64x64 block + FP32 deviation, FP64
 * dt reduction and Kahan-compensated clock), or the vector-mode DualN<N>
 * (autograd/src/dual_n.rs, Real impl real.rs:161-195) for N = 2, 4, 8, 16
 * derivatives in one pass, with the reference's DualN arithmetic (the N
 * the reference's vector-mode example instantiates,
 * solver-2d/examples/m1_forward_scaling.rs:184-187).
 *
 * Tangent arrays at the boundary: Dual, one per value (same layout as the
 * values); DualN<N>, N per value, innermost: the state tangents are
 * [cell][field][k] (n_cells x 3 x N), the bed and Manning tangents
 * [cell][k] (n_cells x N). The batch API takes f64, Dual and FP32. */
enum hf2d_scalar_kind {
  HF2D_F64 = 0,
  HF2D_DUAL = 1,
  HF2D_F32 = 2,
  HF2D_DUAL2 = 3,
  HF2D_DUAL4 = 4,
  HF2D_DUAL8 = 5,
  HF2D_DUAL16 = 6,
  /* FP64 with FMA contraction, reciprocal-and-Newton division (1/h shared
   * by hu/h and hv/h), Newton square root and interior faces evaluated in
   * primitive variables: the f64 path's arithmetic to about one ulp per
   * operation, not bit for bit; no reference counterpart. Single grids
   * only. */
  HF2D_F64_FAST = 7
};

/* Boundary kinds (boundary.rs:80-106). Sides are ordered N, S, W, E
 * (Boundaries2D, boundary.rs:131-140). param is q for Discharge, h for
 * Depth, ignored otherwise. */
enum hf2d_bc_kind {
  HF2D_BC_TRANSMISSIVE = 0,
  HF2D_BC_WALL = 1,
  HF2D_BC_DISCHARGE = 2,
  HF2D_BC_DEPTH = 3
};

/* Integrator (sim.rs:37-46). */
enum hf2d_integrator { HF2D_FORWARD_EULER = 0, HF2D_SSPRK2 = 1 };

/* hf2d_config.options. By default the fused driver (hf2d_step, run_until,
 * run_steps, enqueue_steps) runs each step over a compacted list of active
 * tiles, rebuilt every step on the device: the reference's cell-mask early
 * skip (update.rs:871-900, 1060-1090) at tile granularity. Tiles that have
 * been dry, with a dry neighbourhood, for two steps are left out; their
 * state is a fixed point of the step, so the results are the same bits.
 * Used for f64 and f32 grids without rain and with friction off on dry
 * cells (friction_dry_tol >= 1e-6) of more than 512 tiles; otherwise every
 * tile runs. On fewer tiles every tile runs at once, so leaving some out
 * saves no time and the list's upkeep costs some. HF2D_OPT_NO_ACTIVE_TILES
 * turns the list off (for timing comparisons); HF2D_OPT_ACTIVE_TILES_ANY_SIZE
 * uses it whatever the grid's size (for tests of the list on small grids). */
enum hf2d_option { HF2D_OPT_NO_ACTIVE_TILES = 1, HF2D_OPT_ACTIVE_TILES_ANY_SIZE = 2 };

/* PointSource (source.rs:36-44). */
typedef struct hf2d_point_source {
  int64_t row;
  int64_t col;
  double q_mass; /* m^3/s; negative is a sink */
} hf2d_point_source;

/* SimulationConfig (sim.rs:49-77) plus the forcing the reference's
 * operational callers apply after each step (uk_ea_test8a_official.rs:
 * 206-217): friction dry tolerance (Simulation uses H_DRY, sim.rs:224),
 * uniform rain rate [m/s], and the point sources passed to hf2d_configure. */
typedef struct hf2d_config {
  double cfl;              /* (0, 1]; default 0.4 */
  int32_t integrator;      /* hf2d_integrator; default SSP-RK2 */
  int32_t cfl_with_bcs;    /* 1: cfl_time_step_with_bcs (Simulation); 0: cfl_time_step */
  double max_dt;           /* > 0; default 60 s */
  int32_t friction;        /* apply manning_friction_step after the integrator */
  int32_t options;         /* bit flags, HF2D_OPT_*; 0 = defaults */
  double friction_dry_tol; /* default H_DRY = 1e-6 */
  int64_t max_steps;       /* run_until budget; default 1e7 */
  double rain_rate;        /* apply_rain rate [m/s] after friction; 0 disables */
} hf2d_config;

typedef struct hf2d_sim hf2d_sim;

/* Largest grid side (rows or columns) plus one: cell coordinates are 32-bit
 * on the device. */
#define HF2D_MAX_SIDE (1LL << 30)

/* Mesh2DG::new / StepWorkspace2D::new (geometry.rs:58-76, update.rs:756-771).
 * Fails with HF2D_INVALID_CONFIG on dx <= 0, dy <= 0, an empty grid or a
 * side of HF2D_MAX_SIDE cells or more. */
int hf2d_create(int64_t n_rows, int64_t n_cols, double dx, double dy, int32_t scalar_kind,
                hf2d_sim** out);
int hf2d_destroy(hf2d_sim* sim);
const char* hf2d_last_error(const hf2d_sim* sim);

/* Mesh2DG::with_manning_field (geometry.rs:90-113): bed z and Manning n per
 * cell. Tangent arrays (Dual kind only) may be null (tangent 0). Negative n
 * fails with HF2D_INVALID_CONFIG. */
int hf2d_upload_mesh(hf2d_sim* sim, const double* bed, const double* bed_tangent,
                     const double* manning, const double* manning_tangent);

/* The caller-owned Array2<Conserved2DG<T>> of the reference: copy in / out. */
int hf2d_upload_state(hf2d_sim* sim, const double* state_aos, const double* tangent_aos);
int hf2d_download_state(hf2d_sim* sim, double* state_aos, double* tangent_aos);

/* Boundaries2D by value (boundary.rs:131-140); order N, S, W, E. */
int hf2d_set_boundaries(hf2d_sim* sim, const int32_t kinds[4], const double params[4]);

/* max_wave_speeds (update.rs:113-136). */
int hf2d_max_wave_speeds(hf2d_sim* sim, double* s_x, double* s_y);

/* cfl_time_step (update.rs:148-156) or, with with_bcs != 0,
 * cfl_time_step_with_bcs (update.rs:170-208). +inf for a dry domain. */
int hf2d_cfl_time_step(hf2d_sim* sim, double cfl, int32_t with_bcs, double* dt);

/* forward_euler_step_with (update.rs:805-1211) with caller dt. */
int hf2d_forward_euler_step(hf2d_sim* sim, double dt);

/* ssprk2_step_with (update.rs:1255-1287) with caller dt. */
int hf2d_ssprk2_step(hf2d_sim* sim, double dt);

/* manning_friction_step (source.rs:130-165). */
int hf2d_manning_friction_step(hf2d_sim* sim, double dt, double dry_tol);

/* apply_rain (source.rs:88-101). */
int hf2d_apply_rain(hf2d_sim* sim, double rate, double dt);

/* apply_point_sources (source.rs:52-69), applied in list order. */
int hf2d_apply_point_sources(hf2d_sim* sim, const hf2d_point_source* sources, int64_t n_sources,
                             double dt);

/* Simulation::new's checks and config (sim.rs:128-162): cfl in (0, 1],
 * max_dt > 0, else HF2D_INVALID_CONFIG. Sources are copied. */
int hf2d_configure(hf2d_sim* sim, const hf2d_config* config, const hf2d_point_source* sources,
                   int64_t n_sources);

/* Simulation::step (sim.rs:204-228) followed by the configured rain and
 * point sources. dt = min(CFL, max_dt, dt_cap) is computed on the device;
 * the dt taken is returned. */
int hf2d_step(hf2d_sim* sim, double dt_cap, double* dt_taken);

/* Time-varying forcing of the driver, the operational loops' pattern
 * (uk_ea_test4_official.rs: apply_point_sources(Q(t_new)) after each
 * step; uk_ea_test8a_official.rs:206-217: apply_rain(rate(t_new)) then the
 * point sources): a piecewise-linear series evaluated on the device at
 * t_new = t + dt of each step as
 *     value = (interp(t_new / time_div) * scale_mul) / scale_div
 * with interp as uk_ea_common/mod.rs (clamped at both ends). which =
 * HF2D_SERIES_RAIN replaces the configured rain rate [m/s];
 * HF2D_SERIES_SOURCES replaces the rate q_mass [m^3/s] of every configured
 * point source. n = 0 removes the series. Applies to hf2d_step, run_until,
 * run_steps and enqueue_steps. */
enum hf2d_series_kind { HF2D_SERIES_RAIN = 0, HF2D_SERIES_SOURCES = 1 };
int hf2d_set_forcing_series(hf2d_sim* sim, int32_t which, int32_t n, const double* t, const double* v,
                            double time_div, double scale_mul, double scale_div);

/* Gauges: the operational loops record h at control cells after every
 * step (uk_ea_test4_official.rs, sim_depths). Here the driver appends a row
 * (t, h at each of the n cells) to a device buffer of `capacity` rows after
 * every step; the current (t, h) is the first row. Rows past capacity are
 * dropped (hf2d_read_gauges reports how many were recorded). n = 0 removes
 * the gauges. */
int hf2d_set_gauges(hf2d_sim* sim, int32_t n, const int64_t* rows, const int64_t* cols, int64_t capacity);
/* Copy up to max_rows rows of (1 + n) doubles into out (nullable);
 * *n_rows = rows recorded so far (may exceed capacity if rows were dropped:
 * it is then capped at capacity). */
int hf2d_read_gauges(hf2d_sim* sim, int64_t max_rows, double* out, int64_t* n_rows);

/* Simulation::run_until (sim.rs:231-246). The clock, dt and the step budget
 * stay on the device; the host only polls for completion. */
int hf2d_run_until(hf2d_sim* sim, double t_target);

/* n steps of Simulation::step with no cap (a fixed-count loop). */
int hf2d_run_steps(hf2d_sim* sim, int64_t n_steps);

/* Simulation::time / steps (sim.rs:164-172), and a setter for resuming a
 * saved run. */
int hf2d_get_time(hf2d_sim* sim, double* t, int64_t* steps);
int hf2d_set_time(hf2d_sim* sim, double t, int64_t steps);

/* Measurement: enqueue n device-clocked steps without waiting (no host
 * round trip per step), and per-kernel HIP-event timing on the handle's
 * stream. stage_ms[0] / stage_ms[1] are the summed durations of the first
 * and the final stage kernels over `count` steps since timing was enabled. */
int hf2d_enqueue_steps(hf2d_sim* sim, int64_t n_steps);
int hf2d_synchronize(hf2d_sim* sim);
int hf2d_set_kernel_timing(hf2d_sim* sim, int32_t enabled);
int hf2d_kernel_timing(hf2d_sim* sim, double stage_ms[2], int64_t* count);

/* Total water volume sum(h) dx dy of the current state, summed in FP64 with
 * compensation (the mass budget of the reference's tests, e.g.
 * update.rs:1689-1717, computed on the device). */
int hf2d_total_volume(hf2d_sim* sim, double* volume);

/* hf2d_total_volume without allocating: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the same
 * reduction and the same result bits, with its per-workgroup partial sums in
 * a 16 KiB library-owned device array (one per GPU, kept until the process
 * ends, whatever the grid's size) instead of a buffer allocated per call.
 * Calls on one GPU share that array: calls that overlap on one GPU (two
 * host threads, two handles on one device) race, so serialise them. */
int hf2d_total_volume_noalloc(hf2d_sim* sim, double* volume);

/* Bandwidth reference: average duration of a streaming device copy of n_bytes
 * (read n_bytes, write n_bytes) over `reps` launches. Also the calibration
 * kernel for hardware traffic counters. */
int hf2d_copy_probe(int64_t n_bytes, int32_t reps, double* ms_avg);

/* Arithmetic-contract probe: for each x[i] the device's sqrt(x), x / y[i],
 * pow(x, 4.0/3.0) and pow(x, 4.0/3.0 - 1.0) as used by the step, written to
 * out[4*i + 0..3]. Lets a test check them bit-for-bit against the host. */
int hf2d_math_probe(const double* x, const double* y, int64_t n, double* out);

/* ------------------------------------------------------------------ batches
 * Many independent scenarios (grids of different sizes, each with its own
 * clock, dt, boundaries and forcing) resident in one HBM arena and advanced
 * together: one launch per RK stage over all their tiles (one workgroup per
 * tile), the step captured in a hipGraph. No counterpart in the reference,
 * which runs one Simulation at a time; each scenario's arithmetic is that of
 * hf2d_run_until (sim.rs:231-246) with the configured forcing. */
typedef struct hf2d_batch hf2d_batch;

/* Most scenario slots of one batch (each slot holds 8 KiB of CFL slots). */
#define HF2D_BATCH_MAX_SLOTS 16384

/* HF2D_INVALID_ARGUMENT unless 0 < max_slots <= HF2D_BATCH_MAX_SLOTS. */
int hf2d_batch_create(int32_t scalar_kind, int64_t capacity_bytes, int32_t max_slots, hf2d_batch** out);
int hf2d_batch_destroy(hf2d_batch* batch);
const char* hf2d_batch_last_error(const hf2d_batch* batch);

/* HBM bytes a scenario occupies in a batch of the given kind (admission). */
int hf2d_batch_footprint(int32_t scalar_kind, int64_t n_rows, int64_t n_cols, int64_t n_sources,
                         int64_t* bytes);
int hf2d_batch_free_bytes(const hf2d_batch* batch, int64_t* largest_block, int64_t* total);

/* Admit a scenario that runs until its clock reaches t_end (SSP-RK2; the
 * config's integrator must be HF2D_SSPRK2). HF2D_CAPACITY if no slot or
 * memory is free. manning_tangent / tangent_aos: Dual kind only (nullable). */
int hf2d_batch_add(hf2d_batch* batch, int64_t n_rows, int64_t n_cols, double dx, double dy,
                   const double* bed, const double* manning, const double* manning_tangent,
                   const double* state_aos, const double* tangent_aos, const int32_t bc_kinds[4],
                   const double bc_params[4], const hf2d_config* config,
                   const hf2d_point_source* sources, int64_t n_sources, double t_end,
                   int32_t* slot);
int hf2d_batch_remove(hf2d_batch* batch, int32_t slot);

/* Enqueue n steps of every resident scenario (finished ones are no-ops);
 * use_graph replays a captured 16-step graph. Returns without waiting. */
int hf2d_batch_step(hf2d_batch* batch, int64_t n_steps, int32_t use_graph);
int hf2d_batch_synchronize(hf2d_batch* batch);
/* Whether every step enqueued on the batch has finished (*idle = 1), without
 * waiting: lets one host thread keep several batches busy. */
int hf2d_batch_query(hf2d_batch* batch, int32_t* idle);

/* Per-slot clock and state of slots [0, n_slots): arrays may be null. */
int hf2d_batch_poll(hf2d_batch* batch, int32_t n_slots, double* t, int64_t* steps, int32_t* status,
                    int32_t* done, int32_t* used);
int hf2d_batch_download(hf2d_batch* batch, int32_t slot, double* state_aos, double* tangent_aos);
/* hf2d_batch_download for a slot that has finished (reached t_end or
 * failed) without waiting for the steps still enqueued: a finished slot's
 * later steps are no-ops, so its state is final and can be copied while
 * they run. HF2D_INVALID_ARGUMENT if the slot is still running. */
int hf2d_batch_download_done(hf2d_batch* batch, int32_t slot, double* state_aos, double* tangent_aos);

/* Tile-shape / threads-per-workgroup configurations and the autotuner that
 * times each on the resident mix (states are restored afterwards). A new
 * batch starts on the shape measured fastest on a large flood mix: FP32
 * 32 x 16 x 512, FP64 32 x 12 x 512, Dual 24 x 16 x 1024. */
int hf2d_batch_config_count(const hf2d_batch* batch, int32_t* count);
int hf2d_batch_config_info(const hf2d_batch* batch, int32_t index, int32_t* tile_w, int32_t* tile_h,
                           int32_t* threads);
int hf2d_batch_set_config(hf2d_batch* batch, int32_t index);
int hf2d_batch_autotune(hf2d_batch* batch, int32_t steps, int32_t* best, double* ms_per_step);

#ifdef __cplusplus
}
#endif

#endif /* HYDROFLUX_AMD_H */
