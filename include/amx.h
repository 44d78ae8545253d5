/*
 * amx.h -- C ABI of the MI355X-native per-step MHD update (libamx.so).
 *
 * Plain C: device pointers, sizes, scalar structs and an opaque context. No
 * torch or C++ types cross this boundary. Streams are hipStream_t passed as
 * void* (NULL = the default stream). Every entry point returns an amx_status;
 * no exception crosses the ABI.
 *
 * State layout is the reference's (astronomix/variable_registry/
 * registered_variables.py:116-238): a contiguous float32 array
 * (num_vars, nx, ny[, nz]), row-major, with
 *   2D hydro [rho, vx, vy, p]            2D MHD [rho, vx, vy, p, Bx, By, Bz]
 *   3D hydro [rho, vx, vy, vz, p]        3D MHD [rho, vx, vy, vz, p, Bx, By, Bz]
 * Finite-volume boundaries are periodic, open or reflective per side
 * (amx_fv_config.boundary); the finite-difference path is periodic.
 *
 * The reference has no FFI layer of its own; each entry point below replaces
 * one JAX operator of the reference's finite-volume path, cited per function.
 */
#ifndef AMX_H
#define AMX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum amx_status {
  AMX_OK = 0,
  AMX_ERR_INVALID = 1,       /* bad argument or unsupported configuration */
  AMX_ERR_HIP = 2,           /* a HIP runtime call failed */
  AMX_ERR_NOMEM = 3,         /* device allocation failed */
  AMX_ERR_UNSUPPORTED = 4    /* option outside the implemented path */
} amx_status;

/* Riemann solvers (simulation_config.py:41-46 values). */
#define AMX_HLL 0
#define AMX_HLLC 1
#define AMX_HLLC_LM 2           /* low-Mach HLLC (hll.py:206-221) */
#define AMX_LAX_FRIEDRICHS 3    /* grid-wide alpha per direction (_lax_friedrichs.py:21-78) */
#define AMX_HYBRID_HLLC 4       /* HLLC, HLLC-LM in compressing cells (hll.py:264-326) */
#define AMX_AM_HLLC 5           /* the same with low-Mach dissipation control (hll.py:230-254) */
/* Slope limiters (simulation_config.py:29-34 values). */
#define AMX_MINMOD 0
#define AMX_OSHER 1
#define AMX_DOUBLE_MINMOD 2
#define AMX_SUPERBEE 3
#define AMX_VAN_ALBADA 4
#define AMX_VAN_ALBADA_PP 5     /* van Albada with the positivity scaling of reconstruction.py:168-365 */
/* Ghost-cell boundaries of the finite-volume scheme, per side of an axis
 * (_geometry/boundaries.py:74-357; the reference's own codes are OPEN 0,
 * REFLECTIVE 1, PERIODIC 2, renumbered so that a zeroed config is periodic).
 * Open repeats the edge cell; reflective mirrors the two cells next to the
 * edge and negates the normal velocity of the gas state, and in the magnetic
 * sub-step row axis+1 of B and v (By, vy at x walls; Bz, vz at y walls; none
 * at z walls), as _set_specific_var_along_axis does with var index = axis.
 * Both sides of an axis are periodic or neither is. */
#define AMX_BC_PERIODIC 0
#define AMX_BC_OPEN 1
#define AMX_BC_REFLECTIVE 2
/* Magnetic sub-step integrators (simulation_config.py:88-89 values). */
#define AMX_IMPLICIT_MIDPOINT 0
#define AMX_IMPLICIT_EULER 1

/* Static description of one finite-volume simulation; the fields correspond to
 * SimulationConfig / SimulationParams (option_classes/simulation_config.py:160-342,
 * option_classes/simulation_params.py:11-47). */
typedef struct amx_fv_config {
  int dim;                        /* 2 or 3 (config.dimensionality) */
  int nx, ny, nz;                 /* cells per axis; nz = 1 in 2D */
  int mhd;                        /* config.mhd */
  int riemann;                    /* AMX_HLL ... AMX_AM_HLLC (config.riemann_solver) */
  int magnetic_integrator;        /* config.fv_magnetic_integrator */
  int max_fixed_point_iterations; /* 1000 in the reference (_magnetic_field_update.py:141) */
  float gamma;                    /* params.gamma */
  float C_cfl;                    /* params.C_cfl */
  float fixed_point_tol;          /* 1e-5 for SINGLE_PRECISION (_magnetic_field_update.py:134-139) */
  int limiter;                    /* AMX_MINMOD ... AMX_VAN_ALBADA_PP (config.limiter) */
  double grid_spacing;            /* config.grid_spacing = box_size / num_cells */
  int boundary[6];                /* AMX_BC_* of x low, x high, y low, y high, z low, z high
                                     (config.boundary_settings); all zero = periodic */
} amx_fv_config;

/* Loop controls of time_integration (time_integration.py:629-801). */
typedef struct amx_fv_run {
  float t_end;                    /* params.t_end */
  int fixed_timestep;             /* config.fixed_timestep */
  int num_timesteps;              /* config.num_timesteps */
  int exact_end_time;             /* config.exact_end_time */
  long max_steps;                 /* < 0: no cap (benchmarking uses a cap) */
} amx_fv_run;

typedef struct amx_fv_stats {
  long steps;                     /* steps taken */
  double t;                       /* final simulation time (float32 value) */
  long fixed_point_iterations;    /* sum over steps of while-body iterations */
  int max_fixed_point_iterations; /* largest per-step count */
} amx_fv_stats;

typedef struct amx_fv_ctx amx_fv_ctx;

/* Creates a context on the current HIP device and allocates its workspace:
 * one stage buffer of the gas rows and, with MHD, four sets of fixed-point
 * iterates (B, v) plus the energy-fixed pressure (amx_fv_workspace_bytes).
 * The step loop keeps each step's final B in its iterate set and writes the
 * state's B rows once, at the end of the loop.
 * Replaces the jit/compile setup of time_integration (time_integration.py:108-140). */
amx_status amx_fv_create(const amx_fv_config* cfg, amx_fv_ctx** out);

/* Frees the context and its workspace. */
amx_status amx_fv_destroy(amx_fv_ctx* ctx);

/* Changes gamma and C_cfl of a context (params.gamma / params.C_cfl are
 * traced, not static, in the reference: time_integration.py:108-140), so one
 * context serves simulations that differ only in them. Synchronises the device. */
amx_status amx_fv_set_physics(amx_fv_ctx* ctx, float gamma, float C_cfl);

/* Bytes of device workspace a context of this configuration allocates. */
amx_status amx_fv_workspace_bytes(const amx_fv_config* cfg, size_t* bytes);

/* CFL time step of the unsplit scheme, written to d_dt (one device float).
 * Replaces _cfl_time_step, UNSPLIT branch
 * (_finite_volume/_timestep_estimation/_timestep_estimator.py:94-128). */
amx_status amx_fv_cfl_dt(amx_fv_ctx* ctx, const float* state, float* d_dt, void* stream);

/* One full step in place on `state` with the step size read from the device
 * float d_dt. MHD: S_A(dt/2) o S_B(dt) o S_A(dt/2) with the Pang-Wu fixed
 * point; hydro: one SSP-RK2 step. Blocks the calling thread once per step to
 * read the fixed-point convergence state. *iterations (optional) receives the
 * number of fixed-point while-body iterations.
 * Replaces _evolve_state_fv (_finite_volume/_state_evolution/evolve_state.py:496-600). */
amx_status amx_fv_evolve_state(amx_fv_ctx* ctx, float* state, const float* d_dt, int* iterations,
                               void* stream);

/* Scratch of amx_fv_evolve_state_noalloc: a device array of this many bytes
 * that every GPU the library runs on holds from first use to process exit. */
#define AMX_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)

/* amx_fv_evolve_state without a context and without any device allocation:
 * the same step (same kernels, bit-identical result) with its workspace,
 * control block and constants in the library's static scratch array of
 * AMX_NOALLOC_SCRATCH_BYTES per device instead of a context's allocations.
 * The caller passes the configuration, the state and the device step size.
 * A grid whose workspace (about amx_fv_workspace_bytes(cfg), plus < 1 KiB
 * of alignment) exceeds the scratch is rejected with AMX_ERR_INVALID: 128 MiB
 * holds 3D MHD grids up to about 1.1 M cells (96^3), 2D MHD up to 1.15 M
 * (1024^2), 3D hydro up to 6.7 M. All calls on one device share its scratch
 * and each call re-initialises what it reads; calls that overlap on one
 * device (other streams or threads) race and must not be issued. Like
 * amx_fv_evolve_state it blocks once per MHD step for the fixed point's
 * residuals (read into host stack memory). No counterpart in the reference:
 * it is for callers that cannot allocate device memory -- code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc. */
amx_status amx_fv_evolve_state_noalloc(const amx_fv_config* cfg, float* state, const float* d_dt,
                                       int* iterations, void* stream);

/* Device-resident step loop to t_end (or num_timesteps fixed steps), in place.
 * Replaces the loop of _time_integration (time_integration.py:629-801) with
 * snapshots off. */
amx_status amx_fv_time_integration(amx_fv_ctx* ctx, float* state, const amx_fv_run* run,
                                   amx_fv_stats* stats, void* stream);

/* Reverse mode of one step: lam_in = (dW_out/dW)^T lam_out where W_out is
 * the state amx_fv_evolve_state makes from `state` with the step size in
 * d_dt, held constant (the reference stop_gradients dt,
 * time_integration.py:632-655). `state`, lam_out and lam_in are
 * (num_vars, nx, ny[, nz]) device arrays; `state` is not modified; lam_in
 * must not alias the inputs. The forward pass is recomputed with the forward
 * kernels (the stage states; with mhd = 1 also the magnetic fixed point's
 * iterates, every one kept, and their count). Replaces jax.grad through
 * _evolve_gas_state_unsplit (evolve_state.py:418-488) and, with mhd = 1,
 * through _evolve_state_fv's split step (evolve_state.py:496-600) and the
 * fixed point's checkpointed while loop (_magnetic_field_update.py:171-174),
 * for the differentiation modes of time_integration.py:786-801; the step
 * loop and its checkpointing are the caller's (astronomix_amd.adjoint). HLL /
 * HLLC with minmod, periodic boundaries. */
amx_status amx_fv_adjoint_step(amx_fv_ctx* ctx, const float* state, const float* d_dt, const float* lam_out,
                               float* lam_in, void* stream);

/* Per-kernel-class device time of the launches made while timing is on,
 * measured with HIP events on the launch stream. Fixed-point launches that
 * exit early (already converged) are not counted. */
typedef struct amx_fv_timing {
  double stage_ms;       /* gas stage kernels (k_stage) */
  long stage_launches;
  double mag_ms;         /* magnetic fixed-point kernels (k_mag), executed ones */
  long mag_launches;
} amx_fv_timing;

/* Event timing on ctx: 0 off, 1 every step, n > 1 every n-th step (the
 * events themselves cost time on the stream); resets the counters.
 * No counterpart in the reference (its timing is wall-clock around the jitted
 * loop, time_integration.py:194-234); used by the benchmark harness. */
amx_status amx_fv_set_timing(amx_fv_ctx* ctx, int enable);
amx_status amx_fv_get_timing(amx_fv_ctx* ctx, amx_fv_timing* out);

/* ---------------------------------------------------------------------------
 * Pools: continuous batching of independent finite-volume simulations of one
 * grid and scheme. Every launch of a pool step covers all its running slots
 * (the slot is a grid dimension of each kernel); each slot has its own state,
 * gamma, t_end, dt, CFL reduction and fixed-point iteration count, so its
 * result equals the single-simulation path's on the same input. Slots are
 * loaded and released between steps while the others keep going. No
 * counterpart in the reference (one simulation per time_integration call,
 * time_integration.py:61-236); used by the serving engine.
 * ------------------------------------------------------------------------- */
typedef struct amx_fv_pool amx_fv_pool;

typedef struct amx_fv_slot_status {
  int live;                       /* the slot holds a simulation */
  int finished;                   /* t reached t_end (or num_timesteps steps) */
  long steps;
  double t;
  double t_end;
  long fixed_point_iterations;    /* sum over the slot's steps */
} amx_fv_slot_status;

/* Device bytes one slot of this configuration takes (state + workspace). */
amx_status amx_fv_pool_slot_bytes(const amx_fv_config* cfg, size_t* bytes);
/* A pool of `capacity` empty slots on the current device; `run` gives the
 * step-size rule (adaptive CFL or fixed, exact end); its t_end is unused. */
amx_status amx_fv_pool_create(const amx_fv_config* cfg, const amx_fv_run* run, int capacity,
                              amx_fv_pool** out);
amx_status amx_fv_pool_destroy(amx_fv_pool* pool);
/* Copies a primitive state (host or device memory, reference layout) into a
 * slot and starts it at t = 0 with its own gamma and t_end. */
amx_status amx_fv_pool_load(amx_fv_pool* pool, int slot, const float* state, float gamma, float t_end,
                            void* stream);
/* Copies a slot's state out (host or device memory). With MHD it first
 * writes the slot's magnetic field, which stays in the fixed point's iterate
 * buffers between steps, into the slot's state rows (a kernel on `stream`). */
amx_status amx_fv_pool_read(amx_fv_pool* pool, int slot, float* dst, void* stream);
/* Empties a slot. */
amx_status amx_fv_pool_release(amx_fv_pool* pool, int slot, void* stream);
/* One step of every running slot, in two halves so that several pools (on
 * their own streams) overlap: _launch enqueues the step up to the fixed
 * point's first readback and returns; _finish waits for it, runs any further
 * fixed-point iterations, enqueues the rest of the step and returns. */
amx_status amx_fv_pool_step_launch(amx_fv_pool* pool, void* stream);
amx_status amx_fv_pool_step_finish(amx_fv_pool* pool, void* stream);
/* Pool steps replay HIP graphs captured per (running-slot span, fixed-point
 * batch) by default (env AMX_POOL_GRAPHS=0 turns this off at creation);
 * enable = 0 launches the same kernels directly. */
amx_status amx_fv_pool_set_graphs(amx_fv_pool* pool, int enable);
/* Steps per _launch / _finish (default 1). With steps > 1, _launch queues
 * that many whole steps of every running slot, each step's second half
 * queued behind its fixed-point launches, and the device decides per slot
 * whether those launches settled the fixed point; a slot whose step did not
 * settle does nothing more in the batch, and _finish completes that step with
 * the further launches. One host readback per batch instead of two per step;
 * the results are those of single steps, bit for bit. Slots that finish
 * inside a batch wait for its end. */
amx_status amx_fv_pool_set_batch(amx_fv_pool* pool, int steps);
/* Status of a slot after the last finished step (waits for its readback). */
amx_status amx_fv_pool_status(amx_fv_pool* pool, int slot, amx_fv_slot_status* out);

/* ---------------------------------------------------------------------------
 * Finite-difference MHD (config.solver_mode == FINITE_DIFFERENCE): 3D only,
 * state (11, nx, ny, nz) = [rho, vx, vy, vz, p, Bx, By, Bz, bx, by, bz] where
 * rows 5-7 are cell-centred B and rows 8-10 the face fields at i+1/2 of x, y, z
 * (registered_variables.py:188-198). Characteristic WENO-5 fluxes, HOW-MHD
 * constrained transport, SSPRK(5,4). The run/stats structs are the FV ones.
 * ------------------------------------------------------------------------- */

/* SimulationConfig / SimulationParams fields the FD step reads
 * (simulation_config.py:216, simulation_params.py:29-41). */
typedef struct amx_fd_config {
  int nx, ny, nz;                 /* cells per axis, each >= 6 */
  int enforce_positivity;         /* config.enforce_positivity */
  float gamma;                    /* params.gamma */
  float C_cfl;                    /* params.C_cfl */
  float dt_max;                   /* params.dt_max (0.001 by default) */
  float minimum_density;          /* params.minimum_density */
  float minimum_pressure;         /* params.minimum_pressure */
  float pad_;
  double grid_spacing;            /* config.grid_spacing */
} amx_fd_config;

typedef struct amx_fd_ctx amx_fd_ctx;

/* Creates an FD context on the current HIP device and allocates its
 * workspace (SSPRK registers, interface fluxes, CT edge fields). */
amx_status amx_fd_create(const amx_fd_config* cfg, amx_fd_ctx** out);
amx_status amx_fd_destroy(amx_fd_ctx* ctx);
amx_status amx_fd_workspace_bytes(const amx_fd_config* cfg, size_t* bytes);

/* min(C_cfl * dx / sum_d max|lambda_d|, dt_max) over the 7 MHD wave speeds,
 * written to d_dt. Replaces _cfl_time_step_fd
 * (_finite_difference/_timestep_estimation/_timestep_estimator.py:26-91). */
amx_status amx_fd_cfl_dt(amx_fd_ctx* ctx, const float* state, float* d_dt, void* stream);

/* One SSPRK(5,4) step in place with dt read from the device float d_dt.
 * Replaces _evolve_state_fd (_finite_difference/_state_evolution/_evolve_state.py:25-72). */
amx_status amx_fd_evolve_state(amx_fd_ctx* ctx, float* state, const float* d_dt, void* stream);

/* Device-resident FD step loop, in place (time_integration.py:629-801, FD
 * branch, snapshots off). stats->fixed_point_iterations is always 0. */
amx_status amx_fd_time_integration(amx_fd_ctx* ctx, float* state, const amx_fv_run* run,
                                   amx_fv_stats* stats, void* stream);

/* Face fields from cell-centred B: b_d = (-B[i-1] + 9 B[i] + 9 B[i+1] - B[i+2]) / 16
 * along d, periodic; B and b are (3, nx, ny, nz) and must not alias.
 * Replaces initialize_interface_fields
 * (_finite_difference/_magnetic_update/_constrained_transport.py:201-213). */
amx_status amx_fd_interface_fields(const float* B, float* b, int nx, int ny, int nz, void* stream);

/* Event timing of the WENO flux launches (the dominant kernel class).
 * No counterpart in the reference; used by the benchmark harness. */
amx_status amx_fd_set_timing(amx_fd_ctx* ctx, int enable);
amx_status amx_fd_get_timing(amx_fd_ctx* ctx, double* weno_ms, long* weno_launches);

/* ---------------------------------------------------------------------------
 * Snapshots and diagnostics: the output contract of time_integration with
 * config.return_snapshots (SnapshotData, data_classes/simulation_snapshot_data.py:5-49;
 * time_integration.py:297-369 set-up, :404-537 one snapshot row, :539-610
 * schedule and callbacks, :671-678 dt clamps, :736-745 end-time row).
 * ------------------------------------------------------------------------- */

/* Magnetic-divergence stencils of the max |div B| diagnostic. */
#define AMX_DIV_NONE 0     /* no magnetic field */
#define AMX_DIV_CENTRAL 1  /* cell-centred B, (B[i+1] - B[i-1]) / (2 dx) per axis
                              (divergence3D, _finite_volume/_magnetic_update/_vector_maths.py:33-51) */
#define AMX_DIV_FACES 2    /* face B, sixth-order interface difference
                              (_interface_field_divergence, _finite_difference/_maths/_differencing.py:9-52) */

/* Where the diagnostics find the fields of a state (num_vars, nx, ny[, nz]). */
typedef struct amx_diag_desc {
  int num_vars;                   /* rows of the state */
  int dim;                        /* 2 or 3 */
  int nx, ny, nz;                 /* nz = 1 in 2D */
  int rho_row;                    /* registered_variables.density_index */
  int vel_row;                    /* first of dim velocity rows */
  int p_row;                      /* registered_variables.pressure_index */
  int mag_row;                    /* first of 3 magnetic rows: cell-centred (CENTRAL) or face (FACES) */
  int div_mode;                   /* AMX_DIV_* (CENTRAL needs dim == 3, as divergence3D does) */
  float gamma;                    /* params.gamma */
  double grid_spacing;            /* config.grid_spacing */
  double box_size;                /* config.box_size */
} amx_diag_desc;

/* Snapshot arrays: device pointers to num_snapshots floats each (states:
 * num_snapshots whole states); NULL = the quantity is not recorded (its
 * SnapshotSettings flag is off). The last two fields are outputs. */
typedef struct amx_snapshot_out {
  float* time_points;             /* SnapshotData.time_points */
  float* states;                  /* SnapshotData.states */
  float* total_mass;              /* sum(rho) * box_size^dim (total_quantities.py:155-176) */
  float* total_energy;            /* sum(p/(gamma-1) + rho u^2/2) dx^dim (total_quantities.py:107-150) */
  float* internal_energy;         /* sum(p/(gamma-1)) dx^dim (total_quantities.py:29-43) */
  float* kinetic_energy;          /* sum(rho u^2/2) dx^dim (total_quantities.py:74-85) */
  float* radial_momentum;         /* sum(rho u.r_hat) dx^dim (total_quantities.py:46-71) */
  float* magnetic_divergence;     /* max |div B| (time_integration.py:507-522) */
  int current_checkpoint;         /* out: snapshots taken (SnapshotData.current_checkpoint) */
  long num_iterations;            /* out: steps taken (SnapshotData.num_iterations) */
} amx_snapshot_out;

/* Callback mode (config.activate_snapshot_callback, time_integration.py:571-610):
 * called on the host, with the stream drained, with the simulation time, the
 * checkpoint index and the device state (read only). Non-zero stops the run
 * with AMX_ERR_INVALID. */
typedef int (*amx_snapshot_callback)(void* user, float t, int checkpoint, const float* state);

typedef struct amx_snapshot_spec {
  int num_snapshots;              /* config.num_snapshots */
  int use_timepoints;             /* config.use_specific_snapshot_timepoints */
  const float* timepoints;        /* host array, params.snapshot_timepoints */
  int num_timepoints;             /* its length (>= 1 when use_timepoints) */
  double box_size;                /* config.box_size */
  amx_snapshot_callback callback; /* NULL: record into amx_snapshot_out; else callback mode */
  void* user;                     /* passed to callback */
} amx_snapshot_spec;

/* Bytes of the device workspace amx_snapshot_record needs (block partials). */
size_t amx_snapshot_workspace_bytes(void);

/* One snapshot row (update_snapshot_data, time_integration.py:404-537):
 * out->time_points[index] = t, out->states[index] = state, and each non-NULL
 * diagnostic at [index], for the state described by d. Sums accumulate in
 * float64 in a fixed order (deterministic); the max is exact. work: device
 * buffer of amx_snapshot_workspace_bytes(). Stream-ordered. */
amx_status amx_snapshot_record(const amx_diag_desc* d, const float* state, float t, const amx_snapshot_out* out,
                               int index, void* work, void* stream);

/* The step loops of amx_fv_time_integration / amx_fd_time_integration with
 * the snapshot schedule of time_integration.py:539-610 (a row when
 * t >= k t_end / num_snapshots at the start of a step, or, with timepoints,
 * when |t - timepoints[k]| < 1e-12 at the start of a step or t reaches t_end;
 * dt is then clamped to timepoints[k] - t and exact_end_time is off). Rows
 * past num_snapshots are dropped, as JAX drops out-of-bounds updates. */
amx_status amx_fv_time_integration_snapshots(amx_fv_ctx* ctx, float* state, const amx_fv_run* run,
                                             const amx_snapshot_spec* spec, amx_snapshot_out* out,
                                             amx_fv_stats* stats, void* stream);
amx_status amx_fd_time_integration_snapshots(amx_fd_ctx* ctx, float* state, const amx_fv_run* run,
                                             const amx_snapshot_spec* spec, amx_snapshot_out* out,
                                             amx_fv_stats* stats, void* stream);

/* ---------------------------------------------------------------------------
 * Between-step hooks and turbulent forcing (time_integration.py:694-704;
 * _physics_modules/_turbulent_forcing/_turbulent_forcing.py:14-104).
 * ------------------------------------------------------------------------- */

/* Called by the step loops once per step after dt is set on the device and
 * before the step, on the loop's stream: `state` is the live device state,
 * d_dt the device float holding the step's dt, `step` the step index.
 * Work is enqueued on `stream`; non-zero stops the run with AMX_ERR_INVALID. */
typedef int (*amx_step_hook)(void* user, float* state, const float* d_dt, long step, void* stream);

/* Installs (or, with NULL, removes) the hook of a context's step loop. */
amx_status amx_fv_set_step_hook(amx_fv_ctx* ctx, amx_step_hook hook, void* user);
amx_status amx_fd_set_step_hook(amx_fd_ctx* ctx, amx_step_hook hook, void* user);

/* Forcing spectrum of one step into cw, (3, nx, ny, nz) complex64 (interleaved
 * float pairs): complex normal noise (this library's Philox4x64-10 stream,
 * keyed by seed, counter = (element, step)) times sqrt(P(k)),
 * P(k) = k^6 exp(-8 k / k_p), k_p = 4 pi / box_size, DC removed, compressive
 * part projected out (_turbulent_forcing.py:14-75). The real forcing field is
 * Re(ifftn(cw)) over the three spatial axes (numpy normalisation). */
amx_status amx_forcing_spectrum(float* cw, int nx, int ny, int nz, double box_size, unsigned long long seed,
                                unsigned long long step, void* stream);

/* Bytes of device workspace amx_forcing_apply needs. */
size_t amx_forcing_workspace_bytes(void);

/* Adds amp * w to the velocity rows vel_row..vel_row+2 of `state`, amp the
 * root of a amp^2 + b amp - Edot dt / dx^3 = 0 with a = sum rho |w|^2 / 2,
 * b = sum rho v . w (0 if no real root or |a| <= 1e-10): the kinetic energy
 * grows by energy_injection_rate * dt (_turbulent_forcing.py:77-104). dt is
 * read from the device float d_dt; d_amp (optional) receives amp. */
amx_status amx_forcing_apply(float* state, int num_vars, int nx, int ny, int nz, int rho_row, int vel_row,
                             const float* w, const float* d_dt, float energy_injection_rate, double grid_spacing,
                             void* work, float* d_amp, void* stream);

/* Plain device copy dst[i] = src[i], i < n (16-byte vector loads/stores).
 * No counterpart in the reference: the HBM ceiling (STREAM copy) the
 * benchmark reports next to the kernels, and the known-size kernel against
 * which rocprofv3 FETCH_SIZE / WRITE_SIZE are calibrated. */
amx_status amx_copy_f32(float* dst, const float* src, size_t n, void* stream);

/* Static description of the last error of this thread. */
const char* amx_status_string(amx_status s);
const char* amx_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* AMX_H */
