// Host-side context of one simulation: device buffers (x-fastest SoA, one
// allocation per field), PCG workspace, options and the captured step graph.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "batched.hpp"
#include "device_common.hpp"

struct SmokeNnPreconditioner;
#include "visual-simulation-of-smoke-3d-ext.h"

namespace smoke {

struct HipError : std::runtime_error {
    explicit HipError(const std::string& what) : std::runtime_error(what) {}
};

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

// Scalars of the pressure solve, resident in device memory so a step never
// needs a host round trip in fixed-iteration mode.
struct PcgScalars {
    double rz;          // r . z of the current inner residual
    double alpha;
    double beta;
    double rr;          // inner r . r after the last iteration
    double stop2;       // inner stop threshold, squared
    double bb;          // ||b||^2
    double rr_true;     // ||b - A x||^2 from the last FP64 residual pass
    double tolerance;   // relative tolerance of the whole solve
    double inner_reduction;
    int iteration;      // inner iterations performed in this solve
    int done;           // inner solve finished (tolerance reached or cap)
    int max_iterations;
    int first;          // next direction pass starts a fresh Krylov space (beta = 0)
    int fixed_mode;     // reference semantics: no early exit
    int jacobi;
    double rz_old_new;  // flexible PCG: r_(k+1) . z_k
};

struct PcgStatus {      // what the host polls between chunks
    int iteration;
    int done;
};

// One level of the multilevel preconditioner's hierarchy.  Level 0 borrows
// the pressure grid's occupancy, flags and anchor; its r and z are the
// vectors of the current application.
struct MgLevel {
    int n[3] = {0, 0, 0};
    int64_t cells = 0;
    uint8_t* occ = nullptr;
    uint8_t* flags = nullptr;
    int* anchor = nullptr;
    float* r = nullptr;
    float* z = nullptr;
    float* ta = nullptr;     // sweep results (ping-pong)
    float* tb = nullptr;
    // coarsest level: deflation of v = 1 on the fluid rows
    float* av = nullptr;     // A v
    float* rd = nullptr;     // deflated right-hand side r - c A v
    double* defl = nullptr;  // v^T A v, c = v^T r / v^T A v, then per-block partial sums
};

constexpr int kMaxLevels = 8;

struct Multilevel {
    SmokeSimulationMultilevelOptions options{};
    int allocated = 0;           // levels whose buffers exist
    int built = 0;               // levels whose occupancy / flags match the pressure grid's
    MgLevel level[kMaxLevels];
    double* dot_partials = nullptr;   // current application: per-block r.z of the last sweep (or nullptr)
    int64_t dot_capacity = 0;
    int dot_blocks = 0;
    float* capture = nullptr;    // coarsest-level right-hand sides (training data), not owned
    uint32_t capacity = 0;
    uint32_t captured = 0;
    // kernel timing: a ring of event pairs, each around one run of
    // back-to-back sweeps of the pressure grid (timing_count of them), read
    // back once the ring comes round (by then the pair has long completed) or
    // when the totals are read
    static constexpr int kTimingRing = 64;
    hipEvent_t timing_events[kTimingRing][2] = {};
    bool timing_pending[kTimingRing] = {};
    int timing_count[kTimingRing] = {};
    int timing_next = 0;
};

struct Context {
    SmokeSimulationConfig config{};
    Grid grid{};
    int device = 0;
    hipStream_t stream = nullptr;
    bool owns_stream = false;

    // simulation state
    float* face[3] = {};          // MAC velocity (state)
    float* face_star[3] = {};     // u* (advected, then projected in place of state)
    float* density = nullptr;
    float* temperature = nullptr;
    float* density_source = nullptr;
    float* temperature_source = nullptr;
    float* density_tmp = nullptr;
    float* temperature_tmp = nullptr;
    float* force_user[3] = {};
    float* solid_velocity[3] = {};
    float* solid_temperature = nullptr;
    uint8_t* occupancy = nullptr;
    float* occupancy_float = nullptr;

    // scratch + diagnostics (views)
    float* cell_force[3] = {};
    float* vorticity[3] = {};
    float* vorticity_magnitude = nullptr;
    float* center_velocity[3] = {};
    float* velocity_magnitude = nullptr;
    float* divergence = nullptr;
    float* pressure = nullptr;       // FP32 pressure used by the projection
    float* pressure_rhs = nullptr;

    // occupancy-derived data
    uint8_t* stencil_flags = nullptr;   // bits 0-5 fluid neighbour x-,x+,y-,y+,z-,z+; bit 6 identity row
    int* anchor = nullptr;
    int64_t* solid_pairs = nullptr;     // (solid cell, nearest fluid cell or -1) pairs
    unsigned* solid_count = nullptr;

    // pressure solve workspace
    float* pcg_r = nullptr;
    float* pcg_p = nullptr;            // search direction, updated in place; one plane of zeros follows the cells
    float* pcg_e = nullptr;
    float* pcg_z = nullptr;              // flexible PCG: preconditioned residual (allocated with the network)
    SmokeNnPreconditioner* learned = nullptr;   // attached learned preconditioner, not owned
    double* x64[2] = {};
    int x64_current = 0;
    double* x64_hist = nullptr;          // FP64 solution of the solve before the last one
    int solves_done = 0;                 // tolerance solves since the history was last valid (capped at 2)
    PcgScalars* scalars = nullptr;
    double* partials = nullptr;
    size_t partial_capacity = 0;
    int pcg_stencil_blocks = 0;          // blocks of the last stencil-pass launch (partials to sum)
    int rz_fused_blocks = 0;             // flexible PCG: r.z partials written by the preconditioner (0: none)
    PcgStatus* host_status = nullptr;   // pinned, two slots
    PcgScalars* host_scalars = nullptr; // pinned, two slots (flexible PCG: whole scalar block per iteration)
    hipEvent_t status_event[2] = {};
    // flexible PCG: r.r reduction per iteration of the last inner solve
    // (geometric mean; 0 until one converged), which predicts the end of the
    // next solve before it has a ratio of its own -- scheduling only
    double fpcg_ratio = 0.0;
    // tolerance mode: dt and max |u_face| of the step, copied to pinned memory
    // behind step_event right after the time-step kernel, so that the solver
    // stats need not drain the stream; the anchor read once per occupancy
    float* host_step = nullptr;         // pinned: dt, max |u_face|
    hipEvent_t step_event = nullptr;
    bool step_info_pending = false;
    int anchor_host = -1;               // -1: not read since the occupancy last changed

    // time step
    float* dt_device = nullptr;          // dt used by the current step
    unsigned* max_speed_bits = nullptr;  // running max |u_face| (float bits)
    float* max_speed = nullptr;

    SmokeSimulationSolverOptions options{};
    SmokeSimulationSolverStats stats{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    bool diagnostics_dirty = true;
    // force pass form: the fused kernel (omega staged in LDS) or vorticity_kernel
    // then cell_force_kernel (omega arrays in HBM); the same bits either way
    bool fused_force = true;

    // launch shapes of the tunable kernels: indices into the candidate tables
    // of pcg.hip (stencil passes 2 and 3) and step_kernels.hip (advection of
    // u, v, w and of the scalars); set from the fixed table at creation
    int pcg_shape[2] = {0, 0};
    int advect_block[4] = {0, 0, 0, 0};   // set from the table at creation
    // LDS-staged advection: the outputs (at most one per cell: its tiles
    // cover indices below n) that it leaves to its second pass, and their count
    int* advect_fallback = nullptr;
    unsigned* advect_fallback_count = nullptr;

    Multilevel mg;

    bool timing = false;
    SmokeSimulationKernelTiming timing_totals{};
    hipEvent_t timing_events[2][6] = {};
};

// Batched stages (step_kernels.hip): the launches of the stages other than
// the pressure solve for a batch of contexts, one launch per kernel (and per
// launch form of the plumes' advection kernels), tables in `arena`.
struct BatchedStages {
    std::vector<BatchedLaunch> time_step, forces, advect, project, scalars;
};
void build_batched_stages(const std::vector<Context*>& plumes, TableArena& arena, BatchedStages& out);
void build_batched_sources(const std::vector<Context*>& plumes, const SmokeSimulationPlumeSource* sources, uint32_t step, TableArena& arena,
                           std::vector<BatchedLaunch>& out);

// Launch helpers implemented in the .hip translation units.
void enqueue_fill(Context& c, float* p, float value, int64_t n);
void enqueue_fill_u8(Context& c, uint8_t* p, uint8_t value, int64_t n);
void enqueue_occupancy_derived(Context& c);
void enqueue_time_step(Context& c);
void enqueue_forces(Context& c);
void enqueue_advect_velocity(Context& c);
void enqueue_pressure_rhs(Context& c);
void enqueue_project(Context& c);
void enqueue_scalars(Context& c);
void enqueue_diagnostics(Context& c);
void enqueue_plume_sources(Context& c, const SmokeSimulationPlumeSource& src, uint32_t step);

// Pressure solve (pcg.hip).
void enqueue_pcg_init(Context& c, bool use_x64);          // r = b - A x, workspace reset, reductions
void enqueue_pcg_iteration(Context& c);                    // pass A + pass B
void enqueue_pcg_refine(Context& c);                       // x64 += e, r = b - A x64
void enqueue_pressure_from_x64(Context& c);
void enqueue_fixed_pcg(Context& c);                        // reference: K plain CG iterations from 0
void run_tolerance_pcg(Context& c);                        // host-polled tolerance solve
void enqueue_true_residual(Context& c);                    // rr_true, bb for the FP32 pressure

// Batched solves (pcg.hip): the pressure stage of a batch of contexts in
// Jacobi / plain-CG modes, one launch per pass (and per launch shape of the
// contexts' stencil passes).  Static launches are built once per batch plan;
// the warm start, residual and pressure-from-x64 launches name the contexts'
// current FP64 buffers and are built per solve phase (the warm start rotates
// the buffer roles as run_tolerance_pcg does).
struct BatchedSolve {
    std::vector<BatchedLaunch> rhs, iteration, fixed_start, gather;
};
void build_batched_solve(const std::vector<Context*>& plumes, PcgScalars* gathered, TableArena& arena, BatchedSolve& out);
void build_batched_warm_start(const std::vector<Context*>& plumes, TableArena& arena, std::vector<BatchedLaunch>& out);
void build_batched_residual(const std::vector<Context*>& plumes, bool refine, TableArena& arena, std::vector<BatchedLaunch>& out);
void build_batched_zero_x64(const std::vector<Context*>& plumes, TableArena& arena, std::vector<BatchedLaunch>& out);
void build_batched_pressure_from_x64(const std::vector<Context*>& plumes, TableArena& arena, std::vector<BatchedLaunch>& out);
void configure_pcg(Context& c);
double measure_copy(uint64_t bytes, int reps, double* ms_per_copy);

// Launch-shape table and tuner (pcg.hip, step_kernels.hip).  The tuners time
// candidates on the context's own grid and return the average microseconds
// per launch of the table shape and of the chosen one.
void table_pcg_shapes(Context& c);
void describe_pcg_shape(int index, SmokeSimulationLaunchShape& out);
int find_pcg_shape(const Context& c, const SmokeSimulationLaunchShape& shape);   // a candidate of this grid; throws when none matches
int tune_pcg_shapes(Context& c, SmokeSimulationLaunchReport& report);   // returns candidates timed
void table_advect_blocks(Context& c);

// Multilevel preconditioner (multigrid.hip).
SmokeSimulationMultilevelOptions default_multilevel_options();
bool multilevel_options_valid(const Context& c, const SmokeSimulationMultilevelOptions& o);
void multilevel_build(Context& c);                          // hierarchy of the current occupancy (if stale)
// z = M(r), one V-cycle; with dot_partials, the pressure grid's last sweep
// also writes per-block partial sums of r.z there when it can (returns their
// count, 0 when not fused)
int multilevel_apply(Context& c, const float* r, float* z, double* dot_partials = nullptr, size_t dot_capacity = 0);
void multilevel_release(Context& c);
SmokeSimulationMultilevelTiming multilevel_timing(Context& c, int reps);
void multilevel_drain_timing(Context& c);                  // fold pending sweep events into timing_totals

// Training-pair crops of the last pressure solve (training_data.hip).
void write_training_crops(Context& c, const SmokeSimulationTrainingCrops& crops);
void copy_training_crops(Context& c, const SmokeSimulationDeviceCrops& crops);
void describe_advect_block(int index, SmokeSimulationLaunchShape& out);
int find_advect_block(const SmokeSimulationLaunchShape& shape);   // candidate index; throws when none matches
int tune_advect_blocks(Context& c, SmokeSimulationLaunchReport& report);

}  // namespace smoke

// The ABI's opaque context handle.
struct SmokeSimulationContext_t : smoke::Context {};
