/*
 * Extensions to the smoke-step C ABI (MI355X implementation).
 *
 * None of these entry points exists in the reference; a caller that uses only
 * visual-simulation-of-smoke-3d.h gets the reference semantics unchanged
 * (fixed-dt step, pressure by `pressure_iterations` plain CG iterations from a
 * zero initial guess, backend-cuda.cu:1788-1824).  The extensions add what the
 * reference ABI cannot express:
 *   - a tolerance-terminated pressure solve (Jacobi PCG with FP64 reductions
 *     and FP64 iterative refinement) and its statistics,
 *   - a CFL-adaptive time step computed on the device,
 *   - import/export of the full MAC state (the reference can neither set nor
 *     read face velocities: its velocity view is cell-centred,
 *     backend-cuda.cu:2243-2248),
 *   - running a sub-range of the step's stages (stage-level parity checks),
 *   - on-disk checkpoint / restore of the state,
 *   - per-kernel device timing and a stream-copy bandwidth probe,
 *   - batches of independent plumes of uneven sizes advanced together.
 * Host-pointer entry points here are synchronous; the rest are stream-ordered.
 */
#ifndef VISUAL_SIMULATION_OF_SMOKE_3D_EXT_H
#define VISUAL_SIMULATION_OF_SMOKE_3D_EXT_H

#include <stddef.h>
#include <stdint.h>

#include "visual-simulation-of-smoke-3d.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum SmokeSimulationPressureMode {
    /* Reference semantics: exactly `pressure_iterations` unpreconditioned CG
       iterations from x = 0, no convergence test (backend-cuda.cu:1800-1823). */
    SMOKE_SIMULATION_PRESSURE_FIXED_ITERATIONS = 0,
    /* Iterate until the true residual ||b - A x||_2 / ||b||_2, evaluated in
       FP64, is at most `relative_tolerance`. */
    SMOKE_SIMULATION_PRESSURE_TOLERANCE = 1,
} SmokeSimulationPressureMode;

typedef enum SmokeSimulationPreconditioner {
    SMOKE_SIMULATION_PRECONDITIONER_NONE = 0,
    SMOKE_SIMULATION_PRECONDITIONER_JACOBI = 1,
    /* flexible PCG (Polak-Ribiere beta) with the learned network attached by
       smoke_simulation_ext_set_learned_preconditioner (visual-simulation-of-
       smoke-3d-nn.h); when the inner residual drops by less than 10 % over
       5 iterations, or turns non-finite, the rest of that solve runs
       Jacobi PCG.  The host checks convergence after every iteration. */
    SMOKE_SIMULATION_PRECONDITIONER_LEARNED = 2,
    /* flexible PCG (Polak-Ribiere beta) preconditioned by one V-cycle of a
       geometric hierarchy (smoke_simulation_ext_set_multilevel_options):
       damped-Jacobi smoothing on every level but the coarsest, and at the
       coarsest level either the learned network attached by
       smoke_simulation_ext_set_learned_preconditioner (the network is the
       coarse-grid solver) or damped-Jacobi sweeps only.  FP64 refinement as
       for SMOKE_SIMULATION_PRECONDITIONER_LEARNED; the solve falls back to
       Jacobi PCG when, after 20 iterations, the smallest ||r|| of the last 10
       is not 10 % below the smallest before them, or on a non-finite value
       (||r|| is not monotone under this preconditioner: it rises over the
       first iterations while the near-null mode of the anchored operator is
       built). */
    SMOKE_SIMULATION_PRECONDITIONER_MULTILEVEL = 3,
} SmokeSimulationPreconditioner;

typedef struct SmokeSimulationSolverOptions {
    uint32_t pressure_mode;       /* SmokeSimulationPressureMode */
    uint32_t preconditioner;      /* SmokeSimulationPreconditioner (tolerance mode) */
    double relative_tolerance;    /* e.g. 1e-6 */
    int32_t max_iterations;       /* inner PCG iterations per solve, all passes */
    int32_t max_refinements;      /* FP64 refinement passes after the first inner solve */
    uint32_t warm_start;          /* initial guess: 0 = zero, 1 = the previous step's FP64 pressure,
                                     2 = linear extrapolation 2 x_n - x_(n-1) of the last two solutions
                                     (1 until two exist; the history restarts when the occupancy or the
                                     imported pressure changes) */
    uint32_t adaptive_dt;         /* dt = min(dt_max, cfl * h / max|u_face|) */
    float cfl;
    float dt_max;
    int32_t check_interval;       /* PCG iterations between host convergence polls */
    int32_t reserved;
} SmokeSimulationSolverOptions;

typedef struct SmokeSimulationSolverStats {
    int32_t iterations;           /* inner PCG iterations of the last solve */
    int32_t refinements;          /* FP64 refinement passes of the last solve */
    int32_t converged;            /* tolerance reached (tolerance mode) */
    int32_t anchor;               /* pinned pressure cell (first fluid cell) */
    double relative_residual;     /* true FP64 ||b - A x|| / ||b|| after the last solve */
    double rhs_norm;              /* ||b||_2 */
    float dt;                     /* time step used by the last step */
    float max_face_speed;         /* max |u_face| at the start of the last step */
    uint64_t steps;               /* steps taken by this context */
    int32_t learned_iterations;   /* inner iterations of the last solve preconditioned by the network */
    int32_t learned_fallback;     /* the last solve fell back to Jacobi PCG */
} SmokeSimulationSolverStats;

/* Host-side copy of the simulation state.  Any pointer may be NULL (skipped).
   Face arrays are x-fastest: velocity_x is (nx+1)*ny*nz, velocity_y is
   nx*(ny+1)*nz, velocity_z is nx*ny*(nz+1); cell arrays are nx*ny*nz. */
typedef struct SmokeSimulationHostState {
    float* velocity_x;
    float* velocity_y;
    float* velocity_z;
    float* density;
    float* temperature;
    float* pressure;
    /* intermediate (pre-projection) face velocities, export only */
    float* velocity_star_x;
    float* velocity_star_y;
    float* velocity_star_z;
} SmokeSimulationHostState;

/* Stages of one step, in order. */
typedef enum SmokeSimulationStage {
    SMOKE_SIMULATION_STAGE_FORCES = 0,          /* dt, solid T, vorticity, buoyancy, confinement, faces += dt f */
    SMOKE_SIMULATION_STAGE_ADVECT_VELOCITY = 1, /* u* = advect(u) */
    SMOKE_SIMULATION_STAGE_PRESSURE = 2,        /* b = rhs(u*), solve A p = b */
    SMOKE_SIMULATION_STAGE_PROJECT = 3,         /* u = u* - dt grad p / h */
    SMOKE_SIMULATION_STAGE_SCALARS = 4,         /* T, rho advection, solid fill */
} SmokeSimulationStage;

typedef struct SmokeSimulationKernelTiming {
    uint64_t pcg_iterations;      /* PCG iterations run while timing was on */
    uint64_t pcg_timed_iterations;/* iterations whose three passes were bracketed by events */
    double pcg_direction_ms;      /* summed device time of pass 1: p = z + beta p */
    double pcg_curvature_ms;      /* summed device time of pass 2: p.Ap, alpha */
    double pcg_update_ms;         /* summed device time of pass 3: x += alpha p, r -= alpha Ap, r.z, r.r */
    double pcg_reduction_ms;      /* summed device time of the two one-block finalize kernels (alpha; beta, stop test) */
    double step_ms;               /* summed device time of whole steps */
    uint64_t steps;
    double mg_sweep_ms;           /* multilevel preconditioner: summed device time of the damped-Jacobi
                                     sweeps on the pressure grid (all but the first pre-sweep, which is
                                     x = omega D^-1 r), one event pair around each run of back-to-back
                                     sweeps (the pre-sweeps, the post-sweeps) */
    uint64_t mg_sweeps;           /* sweeps so timed */
} SmokeSimulationKernelTiming;

SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_default_solver_options(SmokeSimulationSolverOptions* out_options);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_set_solver_options(SmokeSimulationContext context, const SmokeSimulationSolverOptions* options);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_get_solver_options(SmokeSimulationContext context, SmokeSimulationSolverOptions* out_options);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_get_solver_stats(SmokeSimulationContext context, SmokeSimulationSolverStats* out_stats);

SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_import_state(SmokeSimulationContext context, const SmokeSimulationHostState* state);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_export_state(SmokeSimulationContext context, const SmokeSimulationHostState* state);
/* The FP64 pressure accumulator of the tolerance-mode solve (nx*ny*nz doubles). */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_export_pressure_f64(SmokeSimulationContext context, double* values);

/* Runs stages first..last (inclusive) of one step; the full step is 0..4. */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_run_stages(SmokeSimulationContext context, int32_t first_stage, int32_t last_stage);

/* Checkpoint file ("SMOKECK1", little-endian, no padding):
     char[8] "SMOKECK1"; u32 version = 2; i32 nx, ny, nz; u32 field count = 18;
     18 times: char[24] name (NUL-padded), u64 byte count, the field's bytes;
     i32 warm-start history count (0..2); end of file.
   Fields, in this order (x-fastest, the state layout): velocity_x/_y/_z
   (MAC faces, FP32), density, temperature, density_source,
   temperature_source, force_x/_y/_z, solid_velocity_x/_y/_z,
   solid_temperature (FP32 cells), occupancy (u8 cells), pressure (FP32),
   pressure_f64, pressure_f64_history (FP64 cells).  save writes the context's
   state; load reads and checks the whole file before touching the context
   (a truncated or mismatched file returns BACKEND_FAILURE and changes
   nothing). */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_save_checkpoint(SmokeSimulationContext context, const char* path);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_load_checkpoint(SmokeSimulationContext context, const char* path);

typedef struct SmokeSimulationCheckpointInfo {
    uint32_t version;
    int32_t dims[3];
    uint32_t field_count;
    int32_t solves_done;
    char error[96];   /* why the file was rejected (empty when accepted) */
} SmokeSimulationCheckpointInfo;

/* Validates a checkpoint file on the host (no device call, no context):
   layout, field names and sizes for its dimensions, end of file. */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_inspect_checkpoint(const char* path, SmokeSimulationCheckpointInfo* out_info);

SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_set_kernel_timing(SmokeSimulationContext context, int32_t enable);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_get_kernel_timing(SmokeSimulationContext context, SmokeSimulationKernelTiming* out_timing);

/* Spherical plume emitter written straight into the density and temperature
   source buffers (the device-side form of the source set-up the reference's
   benchmark does on the host, benchmark.cu:130-147).  Cells whose centre lies
   within `radius` of `center` (both in cell units) get density source
   `density_rate` and temperature source `temperature_rate + temperature_noise
   * U`, U uniform in [-1, 1) from an integer hash of (seed, step, cell); every
   other cell gets 0.  Stream-ordered. */
typedef struct SmokeSimulationPlumeSource {
    float center_x;
    float center_y;
    float center_z;
    float radius;
    float density_rate;
    float temperature_rate;
    float temperature_noise;
    uint32_t seed;
} SmokeSimulationPlumeSource;

SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_plume_sources(SmokeSimulationContext context, const SmokeSimulationPlumeSource* source, uint32_t step);

/* Device-to-device copy of `bytes` bytes, `repetitions` times, on a private
   stream; reports the achieved read+write bandwidth (the stream-copy ceiling
   that roofline fractions are read against). */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_measure_copy_bandwidth(uint64_t bytes, int32_t repetitions, double* out_gb_per_s, double* out_ms_per_copy);

/* Launch shapes.  Replaces the reference's occupancy-derived block choice
   (choose_block, backend-cuda.cu:1881-1911).  By default every kernel uses a
   fixed table chosen from measurements on MI355X, which keeps results
   bitwise reproducible across contexts.  autotune_launches times candidate
   shapes on this context's own grid and state -- the tile of each of the two
   PCG stencil passes and the form and block of each advection kernel --
   applies the fastest and caches the choice for the grid class (dimensions
   and boundary kinds) for the rest of the process; later calls for the same
   class reuse it without timing.  It is synchronous, leaves the simulation
   state unchanged (it writes only the intermediate advected velocity and
   solver scratch, which the next step recomputes) and rebuilds the step
   graph.  A PCG tile other than the table's changes the grouping of the FP64
   partial sums, so results may differ from the table's in the last bits.
   reset_launches returns to the table; set_launches applies the shapes of a
   report (from get_launches or autotune_launches, possibly of another
   process) and fails on a shape that is not a candidate for this grid.

   Advection has two forms with bitwise-equal results.  The gather form runs
   one thread per output and reads every tap from global memory (cells_per_thread
   1, any candidate block).  The LDS-staged form (32 x 8 x 2 threads,
   cells_per_thread 2: tiles of 32 x 8 x 4 outputs, the table's choice for all
   four kernels) stages
   each tile's sources, velocities and occupancy with their ghost values once,
   advects every output whose backtrace (shorter than a cell, no periodic
   wrap) stays inside the staged boxes from LDS, and advects the rest in a
   second pass by the gather form's code. */
typedef struct SmokeSimulationLaunchShape {
    int32_t threads_x;        /* threads along x */
    int32_t threads_y;
    int32_t threads_z;        /* advection: threads along z; PCG: z planes per tile */
    int32_t cells_per_thread; /* PCG: consecutive x cells per thread (4: float4 rows, or 1);
                                 advection: outputs per thread along z (1: gather form, 2: staged form) */
} SmokeSimulationLaunchShape;

typedef struct SmokeSimulationLaunchReport {
    SmokeSimulationLaunchShape pcg_curvature;   /* PCG pass 2 (p.Ap) */
    SmokeSimulationLaunchShape pcg_update;      /* PCG pass 3 (x, r update) */
    SmokeSimulationLaunchShape advect[4];       /* advection of u, v, w, and of T + rho */
    float pcg_curvature_us[2];                  /* [0] table shape, [1] chosen shape (average per launch) */
    float pcg_update_us[2];
    float advect_us[4][2];
    int32_t candidates_timed;                   /* 0 when the grid class was already tuned in this process */
    int32_t from_cache;
} SmokeSimulationLaunchReport;

SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_autotune_launches(SmokeSimulationContext context, SmokeSimulationLaunchReport* out_report);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_reset_launches(SmokeSimulationContext context);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_get_launches(SmokeSimulationContext context, SmokeSimulationLaunchReport* out_report);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_set_launches(SmokeSimulationContext context, const SmokeSimulationLaunchReport* shapes);

/* Form of the force pass (stage 0; no counterpart in the reference, whose
   vorticity and force kernels are B:870-948).  fused != 0 (the default): one
   kernel stages the cell-centred velocity and omega of a marching 64 x 4
   column in LDS and writes the cell force, with no omega arrays in HBM (the
   omega views are recomputed when read).  fused == 0: vorticity_kernel writes
   omega and |omega|, then cell_force_kernel reads them.  Both give the same
   bits.  Rebuilds the step graph. */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_set_force_pass(SmokeSimulationContext context, int32_t fused);

/* Training pairs of the pressure solve, for a learned preconditioner: crops
   of the last solve's right-hand side r = b, its solution z = A^-1 b (the
   FP64 solution in tolerance mode, the FP32 pressure in fixed mode) and the
   operator's per-cell flag byte, packed on the device and appended to a
   binary file.

   File (little-endian): a 64-byte header written when the file is empty --
     char magic[8] = "SMOKETD1"; uint32 version = 1; uint32 crop[3];
     uint32 grid[3]; uint32 record_bytes; uint32 reserved[6] (0)
   -- then one record per crop:
     int32 origin[3]; uint32 step; double scale; uint32 reserved[2];
     uint16 r[N]; uint16 z[N]; uint8 flags[N]      (N = crop[0]*crop[1]*crop[2])
   r and z are bfloat16 (round to nearest even) of b / scale and x / scale,
   x fastest; scale is the RMS of b over the crop (1 when b is zero there),
   so r has unit RMS and A z = r holds with the same scale.  flags: bits 0-5 =
   neighbour x-, x+, y-, y+, z-, z+ is a fluid cell of the system, bit 6 =
   identity row (solid cell or the pressure anchor).  Every crop must lie
   inside the grid.  A line per record is appended to `path` + ".jsonl"
   (offset, origin, step, scale, crop, grid).  Synchronous. */
typedef struct SmokeSimulationTrainingCrops {
    const char* path;
    const int32_t* origins;   /* count x (x, y, z) lower corners */
    uint32_t count;
    uint32_t size[3];         /* crop extent, the same for every record of a file */
    uint32_t step;            /* stored with each record */
} SmokeSimulationTrainingCrops;

SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_write_training_crops(SmokeSimulationContext context, const SmokeSimulationTrainingCrops* crops);

/* The same crops of the last solve's right-hand side, kept on the device for
   training without a file: for crop i, r[i * N + q] = b / scale in FP32 (the
   crop's unit-RMS residual, scale as above, x fastest) and flags[i * N + q]
   the operator's flag byte, N = size[0] * size[1] * size[2].  `origins` is a
   host array of count x (x, y, z); r and flags are device buffers of
   count * N elements.  Ordered on the context's stream (asynchronous). */
typedef struct SmokeSimulationDeviceCrops {
    const int32_t* origins;
    uint32_t count;
    uint32_t size[3];
    float* r;
    uint8_t* flags;
} SmokeSimulationDeviceCrops;

SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_copy_training_crops(SmokeSimulationContext context, const SmokeSimulationDeviceCrops* crops);

/* Attaches (or, with NULL, detaches) a learned preconditioner created by
   smoke_nn_preconditioner_create (an opaque SmokeNnPreconditioner*); the
   context does not own it: keep it alive while attached.  Used when the
   solver options select SMOKE_SIMULATION_PRECONDITIONER_LEARNED. */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_set_learned_preconditioner(SmokeSimulationContext context, void* network);

/* Multilevel preconditioner (SMOKE_SIMULATION_PRECONDITIONER_MULTILEVEL).

   Level 0 is the pressure grid; level l+1 has ceil(n/2) cells per axis of
   level l, and its cell is solid when all of its (up to 8) children are
   solid; the cell holding the previous level's pressure anchor is the
   anchor (an identity row) of the next level.  Every level's operator is the
   7-point matrix of the pressure system rediscretised on that level's
   occupancy (same boundary kinds; a periodic axis must stay even down to the
   coarsest level).  One application z = M(r) is a V-cycle:
     l < L-1: x = `pre_sweeps` damped-Jacobi sweeps from 0;
              r_(l+1) = 1/2 * (sum over the children of r - A x)   (restriction);
              z_(l+1) = M_(l+1)(r_(l+1));
              x += z_(l+1) of the parent cell                   (prolongation);
              z = `post_sweeps` damped-Jacobi sweeps from x;
     l = L-1: z = c v + P^T S(P r), the exact deflation of v = 1 on the
              fluid rows (c = v.r / v.Av, P r = r - c A v,
              P^T y = y - v (Av.y) / v.Av), where S is the network followed by
              `coarse_sweeps` sweeps (coarse_solver = NETWORK) or
              `coarse_sweeps` sweeps from 0 (coarse_solver = JACOBI).
   A damped-Jacobi sweep is x += omega D^-1 (r - A x); identity rows stay 0.
   The deflation is exact for an exact S and takes the near-null mode of the
   anchored operator (nearly constant, set by the global sum of r) out of S's
   task. */
typedef enum SmokeSimulationCoarseSolver {
    SMOKE_SIMULATION_COARSE_NETWORK = 0,
    SMOKE_SIMULATION_COARSE_JACOBI = 1,
} SmokeSimulationCoarseSolver;

typedef struct SmokeSimulationMultilevelOptions {
    int32_t levels;          /* including the pressure grid, >= 2 */
    int32_t pre_sweeps;      /* >= 1 */
    int32_t post_sweeps;     /* >= 1 */
    float omega;             /* damping of the Jacobi sweeps, (0, 1] */
    uint32_t coarse_solver;  /* SmokeSimulationCoarseSolver */
    int32_t coarse_sweeps;   /* >= 0 (NETWORK), >= 1 (JACOBI) */
    int32_t reserved[2];
} SmokeSimulationMultilevelOptions;

/* Defaults: 5 levels (a 16^3 coarsest level for a 256^3 grid), 4 pre- and 4
   post-sweeps, omega 6/7 (the smallest smoothing factor of damped Jacobi on
   the 3D 7-point Laplacian), network coarse solver followed by 8 sweeps. */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_default_multilevel_options(SmokeSimulationMultilevelOptions* out_options);
/* Validates and stores the options; the hierarchy is (re)built on the next
   multilevel solve.  Fails when a periodic axis does not halve evenly down to
   the coarsest level. */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_set_multilevel_options(SmokeSimulationContext context, const SmokeSimulationMultilevelOptions* options);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_get_multilevel_options(SmokeSimulationContext context, SmokeSimulationMultilevelOptions* out_options);

/* Builds the hierarchy for the current occupancy if needed and reports the
   dimensions of each level (levels beyond `levels` are zero). */
typedef struct SmokeSimulationMultilevelInfo {
    int32_t levels;
    int32_t dims[8][3];
    uint32_t captured;       /* deflated coarsest-level right-hand sides stored since the capture was set */
} SmokeSimulationMultilevelInfo;

SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_get_multilevel_info(SmokeSimulationContext context, SmokeSimulationMultilevelInfo* out_info);

/* One application z = M(r) of the V-cycle on device vectors of the pressure
   grid (stream-ordered).  r must be 0 on identity rows; z is 0 there. */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_apply_multilevel(SmokeSimulationContext context, const float* r, float* z);

/* Copies the flag byte of one level (bits as in the training crops) to a
   device buffer of dims[0]*dims[1]*dims[2] bytes (stream-ordered). */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_export_level_flags(SmokeSimulationContext context, int32_t level, uint8_t* flags);

/* Device time of the V-cycle's kernels on the pressure grid, measured with
   HIP events on the context's stream: each kernel is launched `repetitions`
   times back to back on the current hierarchy with the solve's residual as
   input (solver scratch only; the state is not touched) and the average per
   launch is reported, together with one whole V-cycle.  Synchronous. */
typedef struct SmokeSimulationMultilevelTiming {
    float sweep_us;          /* one damped-Jacobi sweep on the pressure grid */
    float restrict_us;       /* residual + restriction to level 1 */
    float scale_us;          /* first sweep from zero (x = omega D^-1 r) */
    float prolong_us;        /* prolongation of the level-1 correction */
    float vcycle_us;         /* one whole application z = M(r) */
    int32_t vector_rows;     /* 1 when the pressure grid runs the float4-row kernels */
    int32_t repetitions;
    int32_t reserved;
} SmokeSimulationMultilevelTiming;

SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_time_multilevel(SmokeSimulationContext context, int32_t repetitions,
                                                                              SmokeSimulationMultilevelTiming* out_timing);

/* Training data of the coarse-grid network: while set, every V-cycle copies
   its deflated coarsest-level right-hand side P r (the input of S) into the
   next free slot of `buffer`
   (device memory, `capacity` volumes of the coarsest level's cells, FP32),
   until it is full.  NULL stops the capture; setting a buffer restarts the
   count. */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_set_multilevel_capture(SmokeSimulationContext context, float* buffer, uint32_t capacity);

/* Batched plumes: many independent simulations of uneven grid sizes advanced
   together (the reference drives one context per scene, scene_plume.cpp:246-392).
   A batch owns `lanes` HIP streams (1..16, 0 for 16: measured fastest for 64
   plumes of 32^3-128^3 on MI355X) and one context per plume, created from
   `plumes[i]` with its stream set to the plume's lane (the desc's own stream
   field is ignored).  Plumes are assigned to lanes longest first: in
   decreasing cell count (ties: lower index first), each to the lane with the
   fewest cells so far (ties: lower lane); within a lane they run in index
   order.  The per-plume contexts take every entry point of the context ABI
   (updates, views, options, learned preconditioner); destroy them only
   through destroy_batch.  A plume's results are bitwise those of a context
   created alone from the same desc and driven the same way. */
typedef struct SmokeSimulationBatch_t* SmokeSimulationBatch;

typedef struct SmokeSimulationBatchCreateDesc {
    const SmokeSimulationContextCreateDesc* plumes;
    uint32_t count;
    int32_t lanes;
} SmokeSimulationBatchCreateDesc;

SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_create_batch(const SmokeSimulationBatchCreateDesc* desc, SmokeSimulationBatch* out_batch);
/* Synchronizes every lane, then destroys the plumes' contexts and the lanes.
   NULL is a no-op. */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_destroy_batch(SmokeSimulationBatch batch);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_batch_context(SmokeSimulationBatch batch, uint32_t index, SmokeSimulationContext* out_context,
                                                                            int32_t* out_lane);
/* One step of every plume; with `sources` (one per plume, or NULL) each
   plume's plume_sources(sources[i], step) comes first.
   When every plume is in FIXED_ITERATIONS mode, or every plume solves to a
   tolerance with Jacobi or no preconditioner (no network, kernel timing off),
   the step is fused: each kernel of the step is one launch over all plumes on
   a stream of the batch (after the lanes' earlier work; the lanes wait for it),
   in fixed mode as one captured graph (the call returns without waiting), in
   tolerance mode with every plume's solve stopping on the device and the
   refinement passes over the plumes still open (the call returns once every
   solve has finished).  Otherwise each plume's own step is enqueued in turn on
   its lane from the calling thread.  Either way a plume's results are bitwise
   those of a context created alone and driven the same way. */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_step_batch(SmokeSimulationBatch batch, const SmokeSimulationPlumeSource* sources,
                                                                         uint32_t step);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_synchronize_batch(SmokeSimulationBatch batch);

/* How the last step_batch ran: fused = 0 (per-plume steps), 1 (fused, fixed
   iterations, one graph), 2 (fused, tolerance solves); the fused step's
   launches (stages other than the solve's iterations, right-hand side
   included) and launches per CG iteration, each over every plume. */
typedef struct SmokeSimulationBatchInfo {
    int32_t fused;
    int32_t graph;
    uint64_t steps;
    uint64_t launches_per_step;
    uint64_t launches_per_iteration;
} SmokeSimulationBatchInfo;

SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_get_batch_info(SmokeSimulationBatch batch, SmokeSimulationBatchInfo* out_info);
/* enable = 0 steps every plume's own context in turn even where the fused
   step applies (for comparison); 1 (the default) fuses where it can. */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_set_batch_fusion(SmokeSimulationBatch batch, int32_t enable);
/* The lane assignment rule above, on the host (no device needed):
   out_lane[i] for plume i of `count` plumes with cells[i] cells. */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_ext_batch_lanes(const uint64_t* cells, uint32_t count, int32_t lanes, int32_t* out_lane);

#ifdef __cplusplus
}
#endif

#endif /* VISUAL_SIMULATION_OF_SMOKE_3D_EXT_H */
