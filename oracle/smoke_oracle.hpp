// CPU oracle of the smoke step -- TEST INFRASTRUCTURE ONLY.
//
// A host restatement of the reference step (backend-cuda.cu, "B" below) used
// as the checker of the HIP implementation.  Only tests/, the smoke() entry of
// __graft_entry__.py and the cpu_baseline leg of bench.py load it; the product
// library never links or calls it.
//
// Every routine is templated on the real type R: R = double is the oracle,
// R = float (with FP64 reductions) is the timed CPU baseline.
//
// Parity status per function (how each one is pinned; tests/test_oracle_*.py):
//   index / wrap / ghost loaders  - brute-force special-case tables (B:121-376)
//   monotonic_cubic_1d            - closed forms: endpoints, linear exactness,
//                                   flat-slope rule, overshoot-free on monotone data
//   samplers (lin / cubic)        - affine fields reproduced exactly (linear),
//                                   integer-shift translation exact (both modes)
//   trace_rk2                     - uniform flow -> exact displacement; wall and
//                                   solid clipping give the closed-form bisection point
//   vorticity / confinement       - solid-body rotation gives omega = 2*Omega,
//                                   linear |omega| ramp gives N = ramp direction
//   buoyancy / force->face        - pointwise closed forms
//   pressure operator             - equals D - W of the fluid-cell adjacency graph
//                                   assembled independently (scipy.sparse), plus
//                                   the n = 1, 2 periodic dedupe cases
//   CG / PCG                      - dense FP64 solve (numpy.linalg) on small grids
//   projection                    - post-projection divergence vanishes on
//                                   non-anchor fluid cells
//   density fill                  - global brute-force nearest search
//   whole step (fixed-K CG)       - parity unpinned at the bit level: cuBLAS /
//                                   cuSPARSE FP32 reduction order cannot be
//                                   reproduced without an NVIDIA GPU
#pragma once

#include <cstdint>

#include "../include/visual-simulation-of-smoke-3d.h"

extern "C" {

typedef struct OracleOptions {
    int32_t pressure_mode;    // 0 = fixed-K plain CG from x = 0, 1 = tolerance PCG
    int32_t preconditioner;   // 0 none, 1 Jacobi (tolerance mode)
    double relative_tolerance;
    int32_t max_iterations;
    int32_t max_refinements;
    int32_t warm_start;
    int32_t adaptive_dt;
    double cfl;
    double dt_max;
    double inner_reduction;   // inner solve stops at this fraction of the outer residual
} OracleOptions;

typedef struct OracleStats {
    int32_t iterations;
    int32_t refinements;
    int32_t converged;
    int32_t anchor;
    double relative_residual;
    double rhs_norm;
    double dt;
    double max_face_speed;
} OracleStats;

// All arrays are FP64, x-fastest, sized from the config.
typedef struct OracleState {
    double* u; double* v; double* w;          // MAC faces
    double* us; double* vs; double* ws;       // pre-projection faces (u*)
    double* density; double* temperature;
    double* density_source; double* temperature_source;
    double* force_user_x; double* force_user_y; double* force_user_z;
    double* solid_velocity_x; double* solid_velocity_y; double* solid_velocity_z;
    uint8_t* occupancy;
    double* solid_temperature;
    double* pressure;                         // FP64 solution (warm-start carrier)
    double* pressure_rhs;
    double* force_x; double* force_y; double* force_z;   // total cell force
    double* center_x; double* center_y; double* center_z;
    double* vorticity_x; double* vorticity_y; double* vorticity_z; double* vorticity_magnitude;
    double* divergence; double* velocity_magnitude;
} OracleState;

typedef struct OraclePlumeSource {
    float center_x, center_y, center_z;   // cell units
    float radius;                         // cell units
    float density_rate;
    float temperature_rate;
    float temperature_noise;              // uniform +-amplitude
    uint32_t seed;
} OraclePlumeSource;

int oracle_run_stages(const SmokeSimulationConfig* cfg, const OracleOptions* opt, OracleState* s, int first, int last, OracleStats* stats);
void oracle_diagnostics(const SmokeSimulationConfig* cfg, OracleState* s);

double oracle_monotonic_cubic(double p0, double p1, double p2, double p3, double t);
// field_kind: 0 density scalar, 1 temperature scalar, 2/3/4 velocity x/y/z
double oracle_sample(const SmokeSimulationConfig* cfg, int field_kind, int cubic, const double* field, double x, double y, double z);
double oracle_load(const SmokeSimulationConfig* cfg, int field_kind, const double* field, int i, int j, int k);
double oracle_load_center_component(const SmokeSimulationConfig* cfg, int component, const double* field, int x, int y, int z);
void oracle_trace(const SmokeSimulationConfig* cfg, const double* u, const double* v, const double* w, const uint8_t* occ, double dt, const double* start, double* out);
int oracle_pressure_anchor(const SmokeSimulationConfig* cfg, const uint8_t* occ);
void oracle_pressure_apply(const SmokeSimulationConfig* cfg, const uint8_t* occ, const double* x, double* y);
void oracle_pressure_dense(const SmokeSimulationConfig* cfg, const uint8_t* occ, double* a);
void oracle_pressure_rhs(const SmokeSimulationConfig* cfg, const uint8_t* occ, const double* us, const double* vs, const double* ws, double dt, double* rhs);
void oracle_nearest_fluid(const SmokeSimulationConfig* cfg, const uint8_t* occ, int64_t* index);
void oracle_plume_sources(const SmokeSimulationConfig* cfg, const OraclePlumeSource* src, uint32_t step, double* density_source, double* temperature_source);
uint32_t oracle_noise_bits(uint32_t seed, uint32_t step, uint32_t cell);

// CPU baseline: FP32 state, FP64 reductions, OpenMP over `threads`.
int oracle_f32_benchmark(const SmokeSimulationConfig* cfg, const OracleOptions* opt, const OraclePlumeSource* src,
                         int warmup, int steps, int threads, double* seconds, OracleStats* stats);
int oracle_f32_step_sample(const SmokeSimulationConfig* cfg, const OracleOptions* opt, const OraclePlumeSource* src, int cg_iterations,
                           int threads, double* t_other, double* t_iteration);
}
