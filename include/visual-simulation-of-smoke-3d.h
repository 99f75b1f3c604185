/*
 * This Source Code Form is subject to the terms of the Mozilla Public
 * License, v. 2.0. If a copy of the MPL was not distributed with this file,
 * You can obtain one at https://mozilla.org/MPL/2.0/.
 *
 * The declarations below (enumerator names and values, struct field order and
 * types, the twelve prototypes) are those of the reference header
 * visual-simulation-of-smoke-3d.h:1-162, which is MPL-2.0 licensed; binary
 * compatibility requires them to be identical, so this header carries the
 * reference's licence.
 *
 * C ABI of the smoke step, MI355X implementation.
 *
 * Binary-compatible drop-in for the interface declared by the reference
 * header visual-simulation-of-smoke-3d.h:1-162: same symbol names (the
 * `_cuda` suffix is part of the ABI even though the implementation is HIP),
 * same enumerator values, same struct layouts (checked by the CPU test
 * suite: sizeof(SmokeSimulationConfig) == 260, sizeof(...CreateDesc) == 280,
 * sizeof(SmokeSimulationView) == 56 on LP64).
 *
 * Semantics (all asynchronous and ordered on the context stream):
 *   - `stream` in the create desc is a hipStream_t; NULL makes the context
 *     create and own a non-blocking stream.
 *   - update_* take DEVICE pointers and copy device-to-device on the context
 *     stream; NULL means "all zero" (solid temperature: "all ambient").
 *   - views alias live internal buffers, valid until the next step/destroy.
 *   - OUT_OF_MEMORY is reported for host allocation failure only; every other
 *     failure, device out-of-memory included, is BACKEND_FAILURE.
 */
#ifndef VISUAL_SIMULATION_OF_SMOKE_3D_H
#define VISUAL_SIMULATION_OF_SMOKE_3D_H

#include <stdint.h>

#if defined(__GNUC__) || defined(__clang__)
#define SMOKE_SIMULATION_API __attribute__((visibility("default")))
#else
#define SMOKE_SIMULATION_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Result codes (reference header :22-26). */
typedef enum SmokeSimulationResult {
    SMOKE_SIMULATION_RESULT_OK = 0,
    SMOKE_SIMULATION_RESULT_OUT_OF_MEMORY = 1,
    SMOKE_SIMULATION_RESULT_BACKEND_FAILURE = 2,
} SmokeSimulationResult;

/* Domain-face type for the flow (reference header :28-33). */
typedef enum SmokeSimulationFlowBoundaryType {
    SMOKE_SIMULATION_FLOW_BOUNDARY_NO_SLIP_WALL = 0,
    SMOKE_SIMULATION_FLOW_BOUNDARY_FREE_SLIP_WALL = 1,
    SMOKE_SIMULATION_FLOW_BOUNDARY_OUTFLOW = 2,
    SMOKE_SIMULATION_FLOW_BOUNDARY_PERIODIC = 3,
} SmokeSimulationFlowBoundaryType;

/* Domain-face type for density / temperature (reference header :35-39). */
typedef enum SmokeSimulationScalarBoundaryType {
    SMOKE_SIMULATION_SCALAR_BOUNDARY_FIXED_VALUE = 0,
    SMOKE_SIMULATION_SCALAR_BOUNDARY_ZERO_FLUX = 1,
    SMOKE_SIMULATION_SCALAR_BOUNDARY_PERIODIC = 2,
} SmokeSimulationScalarBoundaryType;

/* One flow face: type, prescribed wall velocity, outflow pressure (:41-47). */
typedef struct SmokeSimulationFlowBoundaryFaceDesc {
    uint32_t type;
    float velocity_x;
    float velocity_y;
    float velocity_z;
    float pressure;
} SmokeSimulationFlowBoundaryFaceDesc;

typedef struct SmokeSimulationFlowBoundaryConfig {
    SmokeSimulationFlowBoundaryFaceDesc x_minus;
    SmokeSimulationFlowBoundaryFaceDesc x_plus;
    SmokeSimulationFlowBoundaryFaceDesc y_minus;
    SmokeSimulationFlowBoundaryFaceDesc y_plus;
    SmokeSimulationFlowBoundaryFaceDesc z_minus;
    SmokeSimulationFlowBoundaryFaceDesc z_plus;
} SmokeSimulationFlowBoundaryConfig;

/* One scalar face: type and the fixed value (:58-61). */
typedef struct SmokeSimulationScalarBoundaryFaceDesc {
    uint32_t type;
    float value;
} SmokeSimulationScalarBoundaryFaceDesc;

typedef struct SmokeSimulationScalarBoundaryConfig {
    SmokeSimulationScalarBoundaryFaceDesc x_minus;
    SmokeSimulationScalarBoundaryFaceDesc x_plus;
    SmokeSimulationScalarBoundaryFaceDesc y_minus;
    SmokeSimulationScalarBoundaryFaceDesc y_plus;
    SmokeSimulationScalarBoundaryFaceDesc z_minus;
    SmokeSimulationScalarBoundaryFaceDesc z_plus;
} SmokeSimulationScalarBoundaryConfig;

/* Interpolation used by every semi-Lagrangian sample (velocity included). */
typedef enum SmokeSimulationScalarAdvectionMode {
    SMOKE_SIMULATION_SCALAR_ADVECTION_LINEAR = 0,
    SMOKE_SIMULATION_SCALAR_ADVECTION_MONOTONIC_CUBIC = 1,
} SmokeSimulationScalarAdvectionMode;

/* Simulation parameters, immutable after create (reference header :77-92). */
typedef struct SmokeSimulationConfig {
    int32_t nx;
    int32_t ny;
    int32_t nz;
    float cell_size;
    float dt;
    int32_t pressure_iterations;
    float ambient_temperature;
    float buoyancy_density_factor;
    float buoyancy_temperature_factor;
    float vorticity_confinement;
    SmokeSimulationScalarAdvectionMode scalar_advection_mode;
    SmokeSimulationFlowBoundaryConfig flow_boundary;
    SmokeSimulationScalarBoundaryConfig density_boundary;
    SmokeSimulationScalarBoundaryConfig temperature_boundary;
} SmokeSimulationConfig;

typedef struct SmokeSimulationContext_t* SmokeSimulationContext;

typedef struct SmokeSimulationContextCreateDesc {
    SmokeSimulationConfig config;
    void* stream;
    float initial_density;
    float initial_temperature;
} SmokeSimulationContextCreateDesc;

/* Replaces reference header :103-104 / backend-cuda.cu:1833-2083. */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_create_context_cuda(const SmokeSimulationContextCreateDesc* desc, SmokeSimulationContext* out_context);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_destroy_context_cuda(SmokeSimulationContext context);

/* Replaces reference header :106-113 / backend-cuda.cu:2085-2159. */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_update_density_cuda(SmokeSimulationContext context, const float* values);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_update_density_source_cuda(SmokeSimulationContext context, const float* values);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_update_temperature_cuda(SmokeSimulationContext context, const float* values);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_update_temperature_source_cuda(SmokeSimulationContext context, const float* values);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_update_force_cuda(SmokeSimulationContext context, const float* values_x, const float* values_y, const float* values_z);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_update_occupancy_cuda(SmokeSimulationContext context, const uint8_t* values);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_update_solid_velocity_cuda(SmokeSimulationContext context, const float* values_x, const float* values_y, const float* values_z);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_update_solid_temperature_cuda(SmokeSimulationContext context, const float* values);

/* Replaces reference header :114 / backend-cuda.cu:2161-2172. */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_step_cuda(SmokeSimulationContext context);

/* View kinds (reference header :116-132). */
typedef enum SmokeSimulationViewKind {
    SMOKE_SIMULATION_VIEW_DENSITY = 0,
    SMOKE_SIMULATION_VIEW_DENSITY_SOURCE = 1,
    SMOKE_SIMULATION_VIEW_TEMPERATURE = 2,
    SMOKE_SIMULATION_VIEW_TEMPERATURE_SOURCE = 3,
    SMOKE_SIMULATION_VIEW_FORCE = 4,
    SMOKE_SIMULATION_VIEW_SOLID_VELOCITY = 5,
    SMOKE_SIMULATION_VIEW_SOLID_TEMPERATURE = 6,
    SMOKE_SIMULATION_VIEW_FLOW_VELOCITY = 7,
    SMOKE_SIMULATION_VIEW_FLOW_VELOCITY_MAGNITUDE = 8,
    SMOKE_SIMULATION_VIEW_FLOW_PRESSURE = 9,
    SMOKE_SIMULATION_VIEW_FLOW_PRESSURE_RHS = 10,
    SMOKE_SIMULATION_VIEW_FLOW_DIVERGENCE = 11,
    SMOKE_SIMULATION_VIEW_FLOW_VORTICITY = 12,
    SMOKE_SIMULATION_VIEW_FLOW_VORTICITY_MAGNITUDE = 13,
    SMOKE_SIMULATION_VIEW_OCCUPANCY = 14,
} SmokeSimulationViewKind;

typedef enum SmokeSimulationViewLayout {
    SMOKE_SIMULATION_VIEW_LAYOUT_F32_3D = 0,
    SMOKE_SIMULATION_VIEW_LAYOUT_F32_3D_SOA3 = 1,
} SmokeSimulationViewLayout;

typedef struct SmokeSimulationViewRequest {
    uint32_t kind;
    void* consumer_stream;
} SmokeSimulationViewRequest;

typedef struct SmokeSimulationView {
    uint32_t layout;
    int32_t nx;
    int32_t ny;
    int32_t nz;
    uint64_t row_stride_bytes;
    uint64_t slice_stride_bytes;
    const float* data0;
    const float* data1;
    const float* data2;
} SmokeSimulationView;

/* Replaces reference header :156 / backend-cuda.cu:2174-2282. */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_simulation_get_view_cuda(SmokeSimulationContext context, const SmokeSimulationViewRequest* request, SmokeSimulationView* out_view);

#ifdef __cplusplus
}
#endif

#endif /* VISUAL_SIMULATION_OF_SMOKE_3D_H */
