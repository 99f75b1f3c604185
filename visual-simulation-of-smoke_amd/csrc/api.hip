// C ABI of the smoke step (include/visual-simulation-of-smoke-3d.h) and its
// extensions (include/visual-simulation-of-smoke-3d-ext.h).
//
// Mirrors the behaviour of backend-cuda.cu:1829-2284: same validation and
// error codes, stream ownership, update/view semantics.  The step is one
// captured hipGraph in fixed-iteration (reference) mode; the tolerance mode
// runs its stages eagerly and polls the solver between chunks of iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "context.hpp"
#include "visual-simulation-of-smoke-3d-ext.h"


namespace smoke {
namespace {

bool valid_flow_face(const SmokeSimulationFlowBoundaryFaceDesc& f) { return f.type <= 3u; }
bool valid_scalar_face(const SmokeSimulationScalarBoundaryFaceDesc& f) { return f.type <= 2u; }

// B:1835-1863
bool valid_config(const SmokeSimulationConfig& c) {
    if (c.nx <= 0 || c.ny <= 0 || c.nz <= 0 || !(c.cell_size > 0.0f) || !(c.dt > 0.0f) || c.pressure_iterations <= 0) return false;
    if (uint32_t(c.scalar_advection_mode) > 1u) return false;
    const SmokeSimulationFlowBoundaryFaceDesc* ff[6] = {&c.flow_boundary.x_minus, &c.flow_boundary.x_plus, &c.flow_boundary.y_minus,
                                                        &c.flow_boundary.y_plus, &c.flow_boundary.z_minus, &c.flow_boundary.z_plus};
    for (int a = 0; a < 3; ++a) {
        if (!valid_flow_face(*ff[2 * a]) || !valid_flow_face(*ff[2 * a + 1])) return false;
        if ((ff[2 * a]->type == kPeriodic) != (ff[2 * a + 1]->type == kPeriodic)) return false;
    }
    for (const SmokeSimulationScalarBoundaryConfig* s : {&c.density_boundary, &c.temperature_boundary}) {
        const SmokeSimulationScalarBoundaryFaceDesc* sf[6] = {&s->x_minus, &s->x_plus, &s->y_minus, &s->y_plus, &s->z_minus, &s->z_plus};
        for (int a = 0; a < 3; ++a) {
            if (!valid_scalar_face(*sf[2 * a]) || !valid_scalar_face(*sf[2 * a + 1])) return false;
            if ((sf[2 * a]->type == kScalarPeriodic) != (sf[2 * a + 1]->type == kScalarPeriodic)) return false;
        }
    }
    // int32 cell indices in the anchor search (the reference has the same limit, B:1344-1345)
    const int64_t cells = int64_t(c.nx) * c.ny * c.nz;
    return cells < (int64_t(1) << 31) - 1;
}

template <class T>
void alloc(T*& p, size_t count) {
    void* q = nullptr;
    hip_check(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)), "hipMalloc");
    p = static_cast<T*>(q);
}

template <class T>
void release(T*& p) {
    if (p) hipFree(p);
    p = nullptr;
}

// the preconditioned residual of flexible PCG (network or multilevel cycle)
void ensure_pcg_z(Context& c) {
    if (c.pcg_z) return;
    alloc(c.pcg_z, size_t(c.grid.cells));
    enqueue_fill(c, c.pcg_z, 0.0f, c.grid.cells);
}

void destroy_resources(Context& c) {
    if (c.graph_exec) hipGraphExecDestroy(c.graph_exec);
    if (c.graph) hipGraphDestroy(c.graph);
    c.graph_exec = nullptr;
    c.graph = nullptr;
    for (int k = 0; k < 3; ++k) {
        release(c.face[k]);
        release(c.face_star[k]);
        release(c.force_user[k]);
        release(c.solid_velocity[k]);
        release(c.cell_force[k]);
        release(c.vorticity[k]);
        release(c.center_velocity[k]);
    }
    for (float** p : {&c.density, &c.temperature, &c.density_source, &c.temperature_source, &c.density_tmp, &c.temperature_tmp, &c.solid_temperature,
                      &c.occupancy_float, &c.vorticity_magnitude, &c.velocity_magnitude, &c.divergence, &c.pressure, &c.pressure_rhs, &c.pcg_r,
                      &c.pcg_p, &c.pcg_e, &c.dt_device, &c.max_speed})
        release(*p);
    release(c.occupancy);
    release(c.stencil_flags);
    release(c.anchor);
    release(c.solid_pairs);
    release(c.advect_fallback);
    release(c.advect_fallback_count);
    release(c.solid_count);
    release(c.x64[0]);
    release(c.x64[1]);
    release(c.x64_hist);
    release(c.pcg_z);
    release(c.scalars);
    release(c.partials);
    release(c.max_speed_bits);
    multilevel_release(c);
    if (c.host_status) hipHostFree(c.host_status);
    c.host_status = nullptr;
    if (c.host_scalars) hipHostFree(c.host_scalars);
    c.host_scalars = nullptr;
    if (c.host_step) hipHostFree(c.host_step);
    c.host_step = nullptr;
    if (c.step_event) hipEventDestroy(c.step_event), c.step_event = nullptr;
    for (hipEvent_t& e : c.status_event)
        if (e) hipEventDestroy(e), e = nullptr;
    for (auto& row : c.timing_events)
        for (hipEvent_t& e : row)
            if (e) hipEventDestroy(e), e = nullptr;
}

SmokeSimulationSolverOptions default_options(const SmokeSimulationConfig& cfg) {
    SmokeSimulationSolverOptions o{};
    o.pressure_mode = SMOKE_SIMULATION_PRESSURE_FIXED_ITERATIONS;
    o.preconditioner = SMOKE_SIMULATION_PRECONDITIONER_JACOBI;
    o.relative_tolerance = 1.0e-6;
    o.max_iterations = 20000;
    o.max_refinements = 2;
    o.warm_start = 2;
    o.adaptive_dt = 0;
    o.cfl = 0.5f;
    o.dt_max = cfg.dt;
    o.check_interval = 16;
    return o;
}

// Inner solves stop once their residual has fallen by this factor (or below
// half the outer target); FP32 Krylov vectors still track the true residual
// well at this depth, and the FP64 refinement pass fixes the rest.
constexpr double kInnerReduction = 1.0e-4;

void upload_solver_constants(Context& c) {
    PcgScalars s{};
    s.tolerance = c.options.relative_tolerance;
    s.inner_reduction = kInnerReduction;
    s.max_iterations = c.options.max_iterations;
    s.fixed_mode = c.options.pressure_mode == SMOKE_SIMULATION_PRESSURE_FIXED_ITERATIONS;
    s.jacobi = !s.fixed_mode && c.options.preconditioner != SMOKE_SIMULATION_PRECONDITIONER_NONE;
    hip_check(hipStreamSynchronize(c.stream), "sync");
    hip_check(hipMemcpy(c.scalars, &s, sizeof(s), hipMemcpyHostToDevice), "upload solver constants");
}

void enqueue_pressure(Context& c) {
    enqueue_pressure_rhs(c);
    if (c.options.pressure_mode == SMOKE_SIMULATION_PRESSURE_FIXED_ITERATIONS) enqueue_fixed_pcg(c);
    else run_tolerance_pcg(c);
}

void enqueue_stages(Context& c, int first, int last) {
    if (first <= 0 && last >= 0) {
        enqueue_time_step(c);
        enqueue_forces(c);
    }
    if (first <= 1 && last >= 1) enqueue_advect_velocity(c);
    if (first <= 2 && last >= 2) enqueue_pressure(c);
    if (first <= 3 && last >= 3) enqueue_project(c);
    if (first <= 4 && last >= 4) enqueue_scalars(c);
    c.diagnostics_dirty = true;
}

void rebuild_graph(Context& c) {
    if (c.graph_exec) hipGraphExecDestroy(c.graph_exec);
    if (c.graph) hipGraphDestroy(c.graph);
    c.graph_exec = nullptr;
    c.graph = nullptr;
    if (c.options.pressure_mode != SMOKE_SIMULATION_PRESSURE_FIXED_ITERATIONS) return;
    hip_check(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
        enqueue_stages(c, 0, 4);
    } catch (...) {
        hipGraph_t g = nullptr;
        hipStreamEndCapture(c.stream, &g);
        if (g) hipGraphDestroy(g);
        throw;
    }
    hip_check(hipStreamEndCapture(c.stream, &c.graph), "hipStreamEndCapture");
    hip_check(hipGraphInstantiate(&c.graph_exec, c.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
}

void select_device(const Context& c) { hip_check(hipSetDevice(c.device), "hipSetDevice"); }

SmokeSimulationResult guarded_call(Context* c, auto&& fn) {
    if (!c) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    try {
        select_device(*c);
        fn(*c);
        return SMOKE_SIMULATION_RESULT_OK;
    } catch (const std::bad_alloc&) {
        return SMOKE_SIMULATION_RESULT_OUT_OF_MEMORY;
    } catch (...) {
        return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    }
}

// Launch shapes chosen by the tuner, per device and grid class, for the whole
// process: dimensions, boundary kinds and interpolation mode decide both the
// work per kernel and the candidate set.
struct LaunchChoice {
    int pcg_shape[2];
    int advect_block[4];
    SmokeSimulationLaunchReport report;
};

using GridClass = std::tuple<int, int, int, int, uint32_t, uint32_t, uint32_t, int>;

GridClass grid_class(const Context& c) {
    uint32_t flow = 0, dens = 0, temp = 0;
    for (int f = 0; f < 6; ++f) {
        flow = flow * 8u + c.grid.flow_type[f];
        dens = dens * 8u + c.grid.density.type[f];
        temp = temp * 8u + c.grid.temperature.type[f];
    }
    return {c.device, c.grid.n[0], c.grid.n[1], c.grid.n[2], flow, dens, temp, c.grid.cubic};
}

std::mutex& launch_cache_mutex() {
    static std::mutex m;
    return m;
}

std::map<GridClass, LaunchChoice>& launch_cache() {
    static std::map<GridClass, LaunchChoice> cache;
    return cache;
}

void describe_launches(const Context& c, SmokeSimulationLaunchReport& r) {
    describe_pcg_shape(c.pcg_shape[0], r.pcg_curvature);
    describe_pcg_shape(c.pcg_shape[1], r.pcg_update);
    for (int k = 0; k < 4; ++k) describe_advect_block(c.advect_block[k], r.advect[k]);
}

void table_launches(Context& c) {
    table_pcg_shapes(c);
    table_advect_blocks(c);
}

void copy_or_fill(Context& c, float* dst, const float* src, float fill, int64_t n) {
    if (src) hip_check(hipMemcpyAsync(dst, src, size_t(n) * sizeof(float), hipMemcpyDeviceToDevice, c.stream), "hipMemcpyAsync D2D");
    else enqueue_fill(c, dst, fill, n);
}

void step(Context& c) {
    if (c.options.pressure_mode == SMOKE_SIMULATION_PRESSURE_FIXED_ITERATIONS && c.graph_exec) {
        if (c.timing) hip_check(hipEventRecord(c.timing_events[0][0], c.stream), "event");
        hip_check(hipGraphLaunch(c.graph_exec, c.stream), "hipGraphLaunch");
        if (c.timing) {
            hip_check(hipEventRecord(c.timing_events[0][1], c.stream), "event");
            hip_check(hipEventSynchronize(c.timing_events[0][1]), "event sync");
            float ms = 0.f;
            hipEventElapsedTime(&ms, c.timing_events[0][0], c.timing_events[0][1]);
            c.timing_totals.step_ms += ms;
            c.timing_totals.steps += 1;
        }
        c.stats.iterations = c.config.pressure_iterations;
    } else {
        hipEvent_t t0 = nullptr, t1 = nullptr;
        if (c.timing) {
            hip_check(hipEventCreate(&t0), "event");
            hip_check(hipEventCreate(&t1), "event");
            hip_check(hipEventRecord(t0, c.stream), "event");
        }
        enqueue_stages(c, 0, 4);
        if (c.timing) {
            hip_check(hipEventRecord(t1, c.stream), "event");
            hip_check(hipEventSynchronize(t1), "event sync");
            float ms = 0.f;
            hipEventElapsedTime(&ms, t0, t1);
            c.timing_totals.step_ms += ms;
            c.timing_totals.steps += 1;
            hipEventDestroy(t0);
            hipEventDestroy(t1);
        }
    }
    c.stats.steps += 1;
    c.diagnostics_dirty = true;
}

// Checkpoint layout: "SMOKECK1", u32 version 2, i32 nx, ny, nz, u32 field
// count, then per field: char[24] name (NUL-padded), u64 byte count, raw
// little-endian data; after the last field an i32 (solves with a valid
// warm-start history, 0..2) and nothing else.  The field list and sizes are
// fixed by the dimensions (checkpoint_layout).
struct CheckpointField {
    const char* name;
    void* device;
    size_t bytes;
};

constexpr uint32_t kCheckpointVersion = 2;

std::vector<CheckpointField> checkpoint_layout(int nx, int ny, int nz) {
    const size_t cells = size_t(nx) * size_t(ny) * size_t(nz);
    const size_t cb = cells * sizeof(float);
    const size_t fu = size_t(nx + 1) * size_t(ny) * size_t(nz) * sizeof(float);
    const size_t fv = size_t(nx) * size_t(ny + 1) * size_t(nz) * sizeof(float);
    const size_t fw = size_t(nx) * size_t(ny) * size_t(nz + 1) * sizeof(float);
    return {
        {"velocity_x", nullptr, fu},
        {"velocity_y", nullptr, fv},
        {"velocity_z", nullptr, fw},
        {"density", nullptr, cb},
        {"temperature", nullptr, cb},
        {"density_source", nullptr, cb},
        {"temperature_source", nullptr, cb},
        {"force_x", nullptr, cb},
        {"force_y", nullptr, cb},
        {"force_z", nullptr, cb},
        {"solid_velocity_x", nullptr, cb},
        {"solid_velocity_y", nullptr, cb},
        {"solid_velocity_z", nullptr, cb},
        {"solid_temperature", nullptr, cb},
        {"occupancy", nullptr, cells},
        {"pressure", nullptr, cb},
        {"pressure_f64", nullptr, cells * sizeof(double)},
        {"pressure_f64_history", nullptr, cells * sizeof(double)},
    };
}

std::vector<CheckpointField> checkpoint_fields(Context& c) {
    auto f = checkpoint_layout(c.config.nx, c.config.ny, c.config.nz);
    void* dev[] = {c.face[0], c.face[1], c.face[2], c.density, c.temperature, c.density_source, c.temperature_source,
                   c.force_user[0], c.force_user[1], c.force_user[2], c.solid_velocity[0], c.solid_velocity[1], c.solid_velocity[2],
                   c.solid_temperature, c.occupancy, c.pressure, c.x64[c.x64_current], c.x64_hist};
    static_assert(sizeof(dev) / sizeof(dev[0]) == 18);
    for (size_t i = 0; i < f.size(); ++i) f[i].device = dev[i];
    return f;
}

// A whole checkpoint file read and validated on the host (no device call):
// magic, version, dimensions (positive; equal to `want` when given), the
// field list of those dimensions in order with matching names and sizes, the
// warm-start count and the end of the file.  Returns an empty reason on
// success; `data` then holds every field's bytes.
struct CheckpointFile {
    int32_t dims[3] = {0, 0, 0};
    uint32_t version = 0;
    int32_t solves_done = 0;
    std::vector<std::vector<char>> data;
};

std::string read_checkpoint(const char* path, const int32_t* want, CheckpointFile& out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return "cannot open the file";
    struct Closer {
        FILE* f;
        ~Closer() { std::fclose(f); }
    } closer{f};
    char magic[8];
    uint32_t count = 0;
    if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "SMOKECK1", 8) != 0) return "not a SMOKECK1 file";
    if (std::fread(&out.version, 4, 1, f) != 1 || out.version != kCheckpointVersion) return "unsupported version";
    if (std::fread(out.dims, 4, 3, f) != 3 || out.dims[0] <= 0 || out.dims[1] <= 0 || out.dims[2] <= 0) return "bad dimensions";
    if (int64_t(out.dims[0]) * out.dims[1] * out.dims[2] >= (int64_t(1) << 31) - 1) return "bad dimensions";
    if (want && (out.dims[0] != want[0] || out.dims[1] != want[1] || out.dims[2] != want[2])) return "dimensions differ from the context";
    const auto layout = checkpoint_layout(out.dims[0], out.dims[1], out.dims[2]);
    if (std::fread(&count, 4, 1, f) != 1 || count != layout.size()) return "wrong field count";
    out.data.assign(layout.size(), {});
    for (size_t i = 0; i < layout.size(); ++i) {
        char name[24];
        uint64_t bytes = 0;
        if (std::fread(name, 1, 24, f) != 24 || std::fread(&bytes, 8, 1, f) != 1) return "truncated field header";
        char expect[24] = {};
        std::strncpy(expect, layout[i].name, sizeof(expect) - 1);
        if (std::memcmp(name, expect, 24) != 0) return std::string("unexpected field in place of ") + layout[i].name;
        if (bytes != layout[i].bytes) return std::string("wrong size of ") + layout[i].name;
        out.data[i].resize(layout[i].bytes);
        if (std::fread(out.data[i].data(), 1, layout[i].bytes, f) != layout[i].bytes) return std::string("truncated ") + layout[i].name;
    }
    if (std::fread(&out.solves_done, 4, 1, f) != 1 || out.solves_done < 0 || out.solves_done > 2) return "bad warm-start count";
    char extra;
    if (std::fread(&extra, 1, 1, f) != 0) return "trailing bytes";
    return {};
}

}  // namespace

}  // namespace smoke

using smoke::Context;
using smoke::hip_check;

extern "C" {

SmokeSimulationResult smoke_simulation_create_context_cuda(const SmokeSimulationContextCreateDesc* desc, SmokeSimulationContext* out_context) {
    if (desc == nullptr || out_context == nullptr) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    if (!smoke::valid_config(desc->config)) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    *out_context = nullptr;
    std::unique_ptr<SmokeSimulationContext_t> ctx{new (std::nothrow) SmokeSimulationContext_t{}};
    if (!ctx) return SMOKE_SIMULATION_RESULT_OUT_OF_MEMORY;
    Context& c = *ctx;
    try {
        c.config = desc->config;
        c.grid = smoke::make_grid(c.config);
        hip_check(hipGetDevice(&c.device), "hipGetDevice");
        c.stream = static_cast<hipStream_t>(desc->stream);
        if (!c.stream) {
            hip_check(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
            c.owns_stream = true;
        }
        const size_t cells = size_t(c.grid.cells);
        for (int k = 0; k < 3; ++k) {
            smoke::alloc(c.face[k], size_t(c.grid.faces[k]));
            smoke::alloc(c.face_star[k], size_t(c.grid.faces[k]));
            smoke::alloc(c.force_user[k], cells);
            smoke::alloc(c.solid_velocity[k], cells);
            smoke::alloc(c.cell_force[k], cells);
            smoke::alloc(c.vorticity[k], cells);
            smoke::alloc(c.center_velocity[k], cells);
        }
        for (float** p : {&c.density, &c.temperature, &c.density_source, &c.temperature_source, &c.density_tmp, &c.temperature_tmp,
                          &c.solid_temperature, &c.occupancy_float, &c.vorticity_magnitude, &c.velocity_magnitude, &c.divergence, &c.pressure,
                          &c.pressure_rhs, &c.pcg_r, &c.pcg_e})
            smoke::alloc(*p, cells);
        // the search direction is followed by one plane of zeros, read by the
        // stencil passes for neighbours outside the domain
        const int64_t plane = int64_t(c.grid.n[0]) * c.grid.n[1];
        smoke::alloc(c.pcg_p, cells + plane);
        smoke::alloc(c.occupancy, cells);
        smoke::alloc(c.stencil_flags, cells);
        smoke::alloc(c.anchor, 1);
        smoke::alloc(c.solid_pairs, 2 * cells);
        smoke::alloc(c.advect_fallback, cells);
        smoke::alloc(c.advect_fallback_count, 1);
        smoke::alloc(c.solid_count, 1);
        smoke::alloc(c.x64[0], cells);
        smoke::alloc(c.x64[1], cells);
        smoke::alloc(c.x64_hist, cells);
        smoke::alloc(c.scalars, 1);
        smoke::configure_pcg(c);
        smoke::table_launches(c);
        smoke::alloc(c.partials, c.partial_capacity);
        smoke::alloc(c.dt_device, 1);
        smoke::alloc(c.max_speed_bits, 1);
        smoke::alloc(c.max_speed, 1);
        void* pinned = nullptr;
        hip_check(hipHostMalloc(&pinned, 2 * sizeof(smoke::PcgStatus), hipHostMallocDefault), "hipHostMalloc");
        c.host_status = static_cast<smoke::PcgStatus*>(pinned);
        pinned = nullptr;
        hip_check(hipHostMalloc(&pinned, 2 * sizeof(smoke::PcgScalars), hipHostMallocDefault), "hipHostMalloc");
        c.host_scalars = static_cast<smoke::PcgScalars*>(pinned);
        for (hipEvent_t& e : c.status_event) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
        pinned = nullptr;
        hip_check(hipHostMalloc(&pinned, 2 * sizeof(float), hipHostMallocDefault), "hipHostMalloc");
        c.host_step = static_cast<float*>(pinned);
        hip_check(hipEventCreateWithFlags(&c.step_event, hipEventDisableTiming), "event");
        for (auto& row : c.timing_events)
            for (hipEvent_t& e : row) hip_check(hipEventCreate(&e), "event");

        // B:1952-1987 initial values
        for (int k = 0; k < 3; ++k) {
            smoke::enqueue_fill(c, c.face[k], 0.0f, c.grid.faces[k]);
            smoke::enqueue_fill(c, c.face_star[k], 0.0f, c.grid.faces[k]);
            for (float* p : {c.force_user[k], c.solid_velocity[k], c.cell_force[k], c.vorticity[k], c.center_velocity[k]})
                smoke::enqueue_fill(c, p, 0.0f, c.grid.cells);
        }
        for (float* p : {c.density_source, c.temperature_source, c.occupancy_float, c.vorticity_magnitude, c.velocity_magnitude, c.divergence,
                         c.pressure, c.pressure_rhs, c.pcg_r, c.pcg_e})
            smoke::enqueue_fill(c, p, 0.0f, c.grid.cells);
        smoke::enqueue_fill(c, c.pcg_p, 0.0f, c.grid.cells + plane);
        smoke::enqueue_fill(c, c.density, desc->initial_density, c.grid.cells);
        smoke::enqueue_fill(c, c.density_tmp, desc->initial_density, c.grid.cells);
        smoke::enqueue_fill(c, c.temperature, desc->initial_temperature, c.grid.cells);
        smoke::enqueue_fill(c, c.temperature_tmp, desc->initial_temperature, c.grid.cells);
        smoke::enqueue_fill(c, c.solid_temperature, c.config.ambient_temperature, c.grid.cells);
        smoke::enqueue_fill_u8(c, c.occupancy, 0, c.grid.cells);
        hip_check(hipMemsetAsync(c.x64[0], 0, cells * sizeof(double), c.stream), "zero x64");
        hip_check(hipMemsetAsync(c.x64[1], 0, cells * sizeof(double), c.stream), "zero x64");
        hip_check(hipMemsetAsync(c.x64_hist, 0, cells * sizeof(double), c.stream), "zero x64");
        hip_check(hipMemsetAsync(c.max_speed_bits, 0, sizeof(unsigned), c.stream), "zero max");
        smoke::enqueue_fill(c, c.dt_device, c.config.dt, 1);
        smoke::enqueue_fill(c, c.max_speed, 0.0f, 1);
        smoke::enqueue_occupancy_derived(c);
        c.options = smoke::default_options(c.config);
        c.mg.options = smoke::default_multilevel_options();
        smoke::upload_solver_constants(c);
        smoke::rebuild_graph(c);
        hip_check(hipStreamSynchronize(c.stream), "create sync");
        *out_context = ctx.release();
        return SMOKE_SIMULATION_RESULT_OK;
    } catch (const std::bad_alloc&) {
        smoke::destroy_resources(c);
        if (c.owns_stream && c.stream) hipStreamDestroy(c.stream);
        return SMOKE_SIMULATION_RESULT_OUT_OF_MEMORY;
    } catch (...) {
        smoke::destroy_resources(c);
        if (c.owns_stream && c.stream) hipStreamDestroy(c.stream);
        return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    }
}

SmokeSimulationResult smoke_simulation_destroy_context_cuda(SmokeSimulationContext context) {
    if (context == nullptr) return SMOKE_SIMULATION_RESULT_OK;
    Context& c = *context;
    hipSetDevice(c.device);
    if (c.stream) hipStreamSynchronize(c.stream);
    smoke::destroy_resources(c);
    if (c.owns_stream && c.stream) hipStreamDestroy(c.stream);
    delete context;
    return SMOKE_SIMULATION_RESULT_OK;
}

SmokeSimulationResult smoke_simulation_update_density_cuda(SmokeSimulationContext context, const float* values) {
    return smoke::guarded_call(context, [&](Context& c) {
        smoke::copy_or_fill(c, c.density, values, 0.0f, c.grid.cells);
        c.diagnostics_dirty = true;
    });
}

SmokeSimulationResult smoke_simulation_update_density_source_cuda(SmokeSimulationContext context, const float* values) {
    return smoke::guarded_call(context, [&](Context& c) { smoke::copy_or_fill(c, c.density_source, values, 0.0f, c.grid.cells); });
}

SmokeSimulationResult smoke_simulation_update_temperature_cuda(SmokeSimulationContext context, const float* values) {
    return smoke::guarded_call(context, [&](Context& c) { smoke::copy_or_fill(c, c.temperature, values, 0.0f, c.grid.cells); });
}

SmokeSimulationResult smoke_simulation_update_temperature_source_cuda(SmokeSimulationContext context, const float* values) {
    return smoke::guarded_call(context, [&](Context& c) { smoke::copy_or_fill(c, c.temperature_source, values, 0.0f, c.grid.cells); });
}

SmokeSimulationResult smoke_simulation_update_force_cuda(SmokeSimulationContext context, const float* vx, const float* vy, const float* vz) {
    return smoke::guarded_call(context, [&](Context& c) {
        smoke::copy_or_fill(c, c.force_user[0], vx, 0.0f, c.grid.cells);
        smoke::copy_or_fill(c, c.force_user[1], vy, 0.0f, c.grid.cells);
        smoke::copy_or_fill(c, c.force_user[2], vz, 0.0f, c.grid.cells);
    });
}

SmokeSimulationResult smoke_simulation_update_occupancy_cuda(SmokeSimulationContext context, const uint8_t* values) {
    return smoke::guarded_call(context, [&](Context& c) {
        if (values) hip_check(hipMemcpyAsync(c.occupancy, values, size_t(c.grid.cells), hipMemcpyDeviceToDevice, c.stream), "hipMemcpyAsync occupancy");
        else smoke::enqueue_fill_u8(c, c.occupancy, 0, c.grid.cells);
        smoke::enqueue_occupancy_derived(c);
    });
}

SmokeSimulationResult smoke_simulation_update_solid_velocity_cuda(SmokeSimulationContext context, const float* vx, const float* vy, const float* vz) {
    return smoke::guarded_call(context, [&](Context& c) {
        smoke::copy_or_fill(c, c.solid_velocity[0], vx, 0.0f, c.grid.cells);
        smoke::copy_or_fill(c, c.solid_velocity[1], vy, 0.0f, c.grid.cells);
        smoke::copy_or_fill(c, c.solid_velocity[2], vz, 0.0f, c.grid.cells);
    });
}

SmokeSimulationResult smoke_simulation_update_solid_temperature_cuda(SmokeSimulationContext context, const float* values) {
    return smoke::guarded_call(context,
                               [&](Context& c) { smoke::copy_or_fill(c, c.solid_temperature, values, c.config.ambient_temperature, c.grid.cells); });
}

SmokeSimulationResult smoke_simulation_step_cuda(SmokeSimulationContext context) {
    return smoke::guarded_call(context, [&](Context& c) { smoke::step(c); });
}

SmokeSimulationResult smoke_simulation_get_view_cuda(SmokeSimulationContext context, const SmokeSimulationViewRequest* request, SmokeSimulationView* out) {
    if (context == nullptr || request == nullptr || out == nullptr) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    Context& c = *context;
    *out = SmokeSimulationView{};
    out->layout = SMOKE_SIMULATION_VIEW_LAYOUT_F32_3D;
    out->nx = c.config.nx;
    out->ny = c.config.ny;
    out->nz = c.config.nz;
    out->row_stride_bytes = uint64_t(c.config.nx) * sizeof(float);
    out->slice_stride_bytes = uint64_t(c.config.nx) * uint64_t(c.config.ny) * sizeof(float);
    auto soa = [&](float* const f[3]) {
        out->layout = SMOKE_SIMULATION_VIEW_LAYOUT_F32_3D_SOA3;
        out->data0 = f[0];
        out->data1 = f[1];
        out->data2 = f[2];
    };
    bool diagnostic = false;
    switch (request->kind) {
        case SMOKE_SIMULATION_VIEW_DENSITY: out->data0 = c.density; break;
        case SMOKE_SIMULATION_VIEW_DENSITY_SOURCE: out->data0 = c.density_source; break;
        case SMOKE_SIMULATION_VIEW_TEMPERATURE: out->data0 = c.temperature; break;
        case SMOKE_SIMULATION_VIEW_TEMPERATURE_SOURCE: out->data0 = c.temperature_source; break;
        case SMOKE_SIMULATION_VIEW_FORCE: soa(c.force_user); break;
        case SMOKE_SIMULATION_VIEW_SOLID_VELOCITY: soa(c.solid_velocity); break;
        case SMOKE_SIMULATION_VIEW_SOLID_TEMPERATURE: out->data0 = c.solid_temperature; break;
        case SMOKE_SIMULATION_VIEW_FLOW_VELOCITY: soa(c.center_velocity), diagnostic = true; break;
        case SMOKE_SIMULATION_VIEW_FLOW_VELOCITY_MAGNITUDE: out->data0 = c.velocity_magnitude, diagnostic = true; break;
        case SMOKE_SIMULATION_VIEW_FLOW_PRESSURE: out->data0 = c.pressure; break;
        case SMOKE_SIMULATION_VIEW_FLOW_PRESSURE_RHS: out->data0 = c.pressure_rhs; break;
        case SMOKE_SIMULATION_VIEW_FLOW_DIVERGENCE: out->data0 = c.divergence, diagnostic = true; break;
        case SMOKE_SIMULATION_VIEW_FLOW_VORTICITY: soa(c.vorticity), diagnostic = true; break;
        case SMOKE_SIMULATION_VIEW_FLOW_VORTICITY_MAGNITUDE: out->data0 = c.vorticity_magnitude, diagnostic = true; break;
        case SMOKE_SIMULATION_VIEW_OCCUPANCY: out->data0 = c.occupancy_float; break;
        default: return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    }
    return smoke::guarded_call(context, [&](Context& cc) {
        // diagnostics are derived from the current state on demand (stream-ordered)
        if (diagnostic && cc.diagnostics_dirty) smoke::enqueue_diagnostics(cc);
        // B:2189-2205: hand the view to another stream through an event
        hipStream_t consumer = static_cast<hipStream_t>(request->consumer_stream);
        if (consumer && consumer != cc.stream) {
            hipEvent_t ready = nullptr;
            hip_check(hipEventCreateWithFlags(&ready, hipEventDisableTiming), "hipEventCreate");
            hipError_t e1 = hipEventRecord(ready, cc.stream);
            hipError_t e2 = e1 == hipSuccess ? hipStreamWaitEvent(consumer, ready, 0) : e1;
            hipEventDestroy(ready);
            hip_check(e2, "view event");
        }
    });
}

// ---------------------------------------------------------------- extensions

SmokeSimulationResult smoke_simulation_ext_default_solver_options(SmokeSimulationSolverOptions* out) {
    if (!out) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    SmokeSimulationConfig cfg{};
    cfg.dt = 0.01f;
    *out = smoke::default_options(cfg);
    return SMOKE_SIMULATION_RESULT_OK;
}

SmokeSimulationResult smoke_simulation_ext_set_solver_options(SmokeSimulationContext context, const SmokeSimulationSolverOptions* o) {
    if (!o) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    if (o->pressure_mode > 1u || o->preconditioner > 3u || !(o->relative_tolerance > 0.0) || o->max_iterations <= 0 || o->max_refinements < 0 ||
        (o->adaptive_dt && (!(o->cfl > 0.0f) || !(o->dt_max > 0.0f))))
        return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    return smoke::guarded_call(context, [&](Context& c) {
        if (o->preconditioner == SMOKE_SIMULATION_PRECONDITIONER_MULTILEVEL && !smoke::multilevel_options_valid(c, c.mg.options))
            throw smoke::HipError("multilevel options do not fit this grid");
        c.options = *o;
        if (c.options.check_interval <= 0) c.options.check_interval = 16;
        if (c.options.preconditioner == SMOKE_SIMULATION_PRECONDITIONER_MULTILEVEL) smoke::ensure_pcg_z(c);
        smoke::upload_solver_constants(c);
        smoke::rebuild_graph(c);
    });
}

SmokeSimulationResult smoke_simulation_ext_get_solver_options(SmokeSimulationContext context, SmokeSimulationSolverOptions* out) {
    if (!context || !out) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    *out = context->options;
    return SMOKE_SIMULATION_RESULT_OK;
}

SmokeSimulationResult smoke_simulation_ext_get_solver_stats(SmokeSimulationContext context, SmokeSimulationSolverStats* out) {
    if (!out) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    return smoke::guarded_call(context, [&](Context& c) {
        if (c.options.pressure_mode == SMOKE_SIMULATION_PRESSURE_FIXED_ITERATIONS) {
            smoke::enqueue_true_residual(c);
            smoke::PcgScalars s{};
            hip_check(hipMemcpyAsync(&s, c.scalars, sizeof(s), hipMemcpyDeviceToHost, c.stream), "read scalars");
            hip_check(hipStreamSynchronize(c.stream), "sync");
            c.stats.rhs_norm = std::sqrt(s.bb);
            c.stats.relative_residual = s.bb > 0 ? std::sqrt(s.rr_true / s.bb) : 0.0;
            c.stats.converged = 0;
            c.stats.refinements = 0;
        }
        float dt = 0.f, speed = 0.f;
        int anchor = c.anchor_host;
        const bool tolerance = c.options.pressure_mode != SMOKE_SIMULATION_PRESSURE_FIXED_ITERATIONS;
        if (tolerance && c.step_info_pending && anchor >= 0) {
            // the copies made right after the last step's time-step kernel:
            // no wait for the rest of the step
            hip_check(hipEventSynchronize(c.step_event), "step info");
            dt = c.host_step[0];
            speed = c.host_step[1];
        } else {
            hip_check(hipMemcpyAsync(&dt, c.dt_device, sizeof(float), hipMemcpyDeviceToHost, c.stream), "read dt");
            hip_check(hipMemcpyAsync(&speed, c.max_speed, sizeof(float), hipMemcpyDeviceToHost, c.stream), "read speed");
            hip_check(hipMemcpyAsync(&anchor, c.anchor, sizeof(int), hipMemcpyDeviceToHost, c.stream), "read anchor");
            hip_check(hipStreamSynchronize(c.stream), "sync");
            c.anchor_host = anchor;
        }
        c.stats.dt = dt;
        c.stats.max_face_speed = speed;
        c.stats.anchor = anchor;
        *out = c.stats;
    });
}

SmokeSimulationResult smoke_simulation_ext_import_state(SmokeSimulationContext context, const SmokeSimulationHostState* s) {
    if (!s) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    return smoke::guarded_call(context, [&](Context& c) {
        hip_check(hipStreamSynchronize(c.stream), "sync");
        const size_t cb = size_t(c.grid.cells) * sizeof(float);
        float* faces[3] = {s->velocity_x, s->velocity_y, s->velocity_z};
        for (int k = 0; k < 3; ++k)
            if (faces[k]) hip_check(hipMemcpy(c.face[k], faces[k], size_t(c.grid.faces[k]) * sizeof(float), hipMemcpyHostToDevice), "import face");
        if (s->density) hip_check(hipMemcpy(c.density, s->density, cb, hipMemcpyHostToDevice), "import density");
        if (s->temperature) hip_check(hipMemcpy(c.temperature, s->temperature, cb, hipMemcpyHostToDevice), "import temperature");
        if (s->pressure) {
            std::vector<double> p(size_t(c.grid.cells));
            for (size_t i = 0; i < p.size(); ++i) p[i] = s->pressure[i];
            hip_check(hipMemcpy(c.pressure, s->pressure, cb, hipMemcpyHostToDevice), "import pressure");
            hip_check(hipMemcpy(c.x64[c.x64_current], p.data(), p.size() * sizeof(double), hipMemcpyHostToDevice), "import pressure f64");
            smoke::enqueue_occupancy_derived(c);   // re-zeroes identity rows of the FP64 pressure
            hip_check(hipStreamSynchronize(c.stream), "sync");
        }
        c.diagnostics_dirty = true;
    });
}

SmokeSimulationResult smoke_simulation_ext_export_state(SmokeSimulationContext context, const SmokeSimulationHostState* s) {
    if (!s) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    return smoke::guarded_call(context, [&](Context& c) {
        hip_check(hipStreamSynchronize(c.stream), "sync");
        const size_t cb = size_t(c.grid.cells) * sizeof(float);
        float* faces[3] = {s->velocity_x, s->velocity_y, s->velocity_z};
        float* stars[3] = {s->velocity_star_x, s->velocity_star_y, s->velocity_star_z};
        for (int k = 0; k < 3; ++k) {
            if (faces[k]) hip_check(hipMemcpy(faces[k], c.face[k], size_t(c.grid.faces[k]) * sizeof(float), hipMemcpyDeviceToHost), "export face");
            if (stars[k]) hip_check(hipMemcpy(stars[k], c.face_star[k], size_t(c.grid.faces[k]) * sizeof(float), hipMemcpyDeviceToHost), "export star");
        }
        if (s->density) hip_check(hipMemcpy(s->density, c.density, cb, hipMemcpyDeviceToHost), "export density");
        if (s->temperature) hip_check(hipMemcpy(s->temperature, c.temperature, cb, hipMemcpyDeviceToHost), "export temperature");
        if (s->pressure) hip_check(hipMemcpy(s->pressure, c.pressure, cb, hipMemcpyDeviceToHost), "export pressure");
    });
}

SmokeSimulationResult smoke_simulation_ext_export_pressure_f64(SmokeSimulationContext context, double* values) {
    if (!values) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    return smoke::guarded_call(context, [&](Context& c) {
        hip_check(hipStreamSynchronize(c.stream), "sync");
        hip_check(hipMemcpy(values, c.x64[c.x64_current], size_t(c.grid.cells) * sizeof(double), hipMemcpyDeviceToHost), "export pressure f64");
    });
}

SmokeSimulationResult smoke_simulation_ext_run_stages(SmokeSimulationContext context, int32_t first, int32_t last) {
    if (first < 0 || last > 4 || first > last) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    return smoke::guarded_call(context, [&](Context& c) { smoke::enqueue_stages(c, first, last); });
}

SmokeSimulationResult smoke_simulation_ext_save_checkpoint(SmokeSimulationContext context, const char* path) {
    if (!path) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    return smoke::guarded_call(context, [&](Context& c) {
        hip_check(hipStreamSynchronize(c.stream), "sync");
        FILE* f = std::fopen(path, "wb");
        if (!f) throw smoke::HipError("cannot open checkpoint for writing");
        auto fields = smoke::checkpoint_fields(c);
        const uint32_t version = smoke::kCheckpointVersion, count = uint32_t(fields.size());
        const int32_t dims[3] = {c.config.nx, c.config.ny, c.config.nz};
        bool ok = std::fwrite("SMOKECK1", 1, 8, f) == 8 && std::fwrite(&version, 4, 1, f) == 1 && std::fwrite(dims, 4, 3, f) == 3 &&
                  std::fwrite(&count, 4, 1, f) == 1;
        std::vector<char> host;
        for (auto& fd : fields) {
            if (!ok) break;
            char name[24] = {};
            std::strncpy(name, fd.name, sizeof(name) - 1);
            const uint64_t bytes = fd.bytes;
            host.resize(fd.bytes);
            if (hipMemcpy(host.data(), fd.device, fd.bytes, hipMemcpyDeviceToHost) != hipSuccess) ok = false;
            ok = ok && std::fwrite(name, 1, 24, f) == 24 && std::fwrite(&bytes, 8, 1, f) == 1 && std::fwrite(host.data(), 1, fd.bytes, f) == fd.bytes;
        }
        const int32_t solves_done = c.solves_done;   // validity of the warm-start history
        ok = ok && std::fwrite(&solves_done, 4, 1, f) == 1;
        ok = (std::fclose(f) == 0) && ok;
        if (!ok) throw smoke::HipError("checkpoint write failed");
    });
}

SmokeSimulationResult smoke_simulation_ext_load_checkpoint(SmokeSimulationContext context, const char* path) {
    if (!path) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    return smoke::guarded_call(context, [&](Context& c) {
        // the whole file is read and checked before any device buffer
        // changes: a bad file leaves the context as it was
        const int32_t want[3] = {c.config.nx, c.config.ny, c.config.nz};
        smoke::CheckpointFile file;
        const std::string why = smoke::read_checkpoint(path, want, file);
        if (!why.empty()) throw smoke::HipError("checkpoint: " + why);
        hip_check(hipStreamSynchronize(c.stream), "sync");
        auto fields = smoke::checkpoint_fields(c);
        for (size_t i = 0; i < fields.size(); ++i)
            hip_check(hipMemcpy(fields[i].device, file.data[i].data(), fields[i].bytes, hipMemcpyHostToDevice), "checkpoint upload");
        smoke::enqueue_occupancy_derived(c);
        c.solves_done = file.solves_done;
        hip_check(hipStreamSynchronize(c.stream), "sync");
    });
}

SmokeSimulationResult smoke_simulation_ext_inspect_checkpoint(const char* path, SmokeSimulationCheckpointInfo* out_info) {
    if (!path || !out_info) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    try {
        smoke::CheckpointFile file;
        const std::string why = smoke::read_checkpoint(path, nullptr, file);
        *out_info = SmokeSimulationCheckpointInfo{};
        std::strncpy(out_info->error, why.c_str(), sizeof(out_info->error) - 1);
        if (!why.empty()) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
        out_info->version = file.version;
        for (int k = 0; k < 3; ++k) out_info->dims[k] = file.dims[k];
        out_info->field_count = uint32_t(file.data.size());
        out_info->solves_done = file.solves_done;
        return SMOKE_SIMULATION_RESULT_OK;
    } catch (const std::bad_alloc&) {
        return SMOKE_SIMULATION_RESULT_OUT_OF_MEMORY;
    } catch (...) {
        return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    }
}

SmokeSimulationResult smoke_simulation_ext_set_kernel_timing(SmokeSimulationContext context, int32_t enable) {
    return smoke::guarded_call(context, [&](Context& c) {
        smoke::multilevel_drain_timing(c);   // pending pairs belong to the old totals
        c.timing = enable != 0;
        c.timing_totals = SmokeSimulationKernelTiming{};
    });
}

SmokeSimulationResult smoke_simulation_ext_get_kernel_timing(SmokeSimulationContext context, SmokeSimulationKernelTiming* out) {
    if (!out) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    return smoke::guarded_call(context, [&](Context& c) {
        smoke::multilevel_drain_timing(c);
        *out = c.timing_totals;
    });
}

SmokeSimulationResult smoke_simulation_ext_plume_sources(SmokeSimulationContext context, const SmokeSimulationPlumeSource* src, uint32_t step) {
    if (!src) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    return smoke::guarded_call(context, [&](Context& c) { smoke::enqueue_plume_sources(c, *src, step); });
}

SmokeSimulationResult smoke_simulation_ext_autotune_launches(SmokeSimulationContext context, SmokeSimulationLaunchReport* out) {
    return smoke::guarded_call(context, [&](Context& c) {
        const smoke::GridClass key = smoke::grid_class(c);
        SmokeSimulationLaunchReport report{};
        bool cached = false;
        {
            std::lock_guard<std::mutex> lock(smoke::launch_cache_mutex());
            auto it = smoke::launch_cache().find(key);
            if (it != smoke::launch_cache().end()) {
                std::copy(std::begin(it->second.pcg_shape), std::end(it->second.pcg_shape), c.pcg_shape);
                std::copy(std::begin(it->second.advect_block), std::end(it->second.advect_block), c.advect_block);
                report = it->second.report;
                cached = true;
            }
        }
        if (!cached) {
            smoke::table_launches(c);
            hip_check(hipStreamSynchronize(c.stream), "sync");
            report.candidates_timed = smoke::tune_advect_blocks(c, report);
            report.candidates_timed += smoke::tune_pcg_shapes(c, report);
            smoke::upload_solver_constants(c);   // the PCG tuner ran with timing scalars
            std::lock_guard<std::mutex> lock(smoke::launch_cache_mutex());
            smoke::LaunchChoice choice{};
            std::copy(std::begin(c.pcg_shape), std::end(c.pcg_shape), choice.pcg_shape);
            std::copy(std::begin(c.advect_block), std::end(c.advect_block), choice.advect_block);
            choice.report = report;
            smoke::launch_cache()[key] = choice;
        }
        smoke::describe_launches(c, report);
        report.from_cache = cached ? 1 : 0;
        if (cached) report.candidates_timed = 0;
        smoke::rebuild_graph(c);
        hip_check(hipStreamSynchronize(c.stream), "sync");
        if (out) *out = report;
    });
}

SmokeSimulationResult smoke_simulation_ext_reset_launches(SmokeSimulationContext context) {
    return smoke::guarded_call(context, [&](Context& c) {
        smoke::table_launches(c);
        smoke::rebuild_graph(c);
    });
}

SmokeSimulationResult smoke_simulation_ext_get_launches(SmokeSimulationContext context, SmokeSimulationLaunchReport* out) {
    if (!out) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    return smoke::guarded_call(context, [&](Context& c) {
        SmokeSimulationLaunchReport r{};
        smoke::describe_launches(c, r);
        *out = r;
    });
}

SmokeSimulationResult smoke_simulation_ext_set_launches(SmokeSimulationContext context, const SmokeSimulationLaunchReport* shapes) {
    if (!shapes) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    return smoke::guarded_call(context, [&](Context& c) {
        // every shape is resolved before any is applied
        const int pcg[2] = {smoke::find_pcg_shape(c, shapes->pcg_curvature), smoke::find_pcg_shape(c, shapes->pcg_update)};
        int advect[4];
        for (int k = 0; k < 4; ++k) advect[k] = smoke::find_advect_block(shapes->advect[k]);
        std::copy(std::begin(pcg), std::end(pcg), c.pcg_shape);
        std::copy(std::begin(advect), std::end(advect), c.advect_block);
        smoke::rebuild_graph(c);
    });
}

SmokeSimulationResult smoke_simulation_ext_set_force_pass(SmokeSimulationContext context, int32_t fused) {
    return smoke::guarded_call(context, [&](Context& c) {
        c.fused_force = fused != 0;
        smoke::rebuild_graph(c);
    });
}

SmokeSimulationResult smoke_simulation_ext_write_training_crops(SmokeSimulationContext context, const SmokeSimulationTrainingCrops* crops) {
    if (!crops) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    return smoke::guarded_call(context, [&](Context& c) { smoke::write_training_crops(c, *crops); });
}

SmokeSimulationResult smoke_simulation_ext_copy_training_crops(SmokeSimulationContext context, const SmokeSimulationDeviceCrops* crops) {
    if (!crops) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    return smoke::guarded_call(context, [&](Context& c) { smoke::copy_training_crops(c, *crops); });
}

SmokeSimulationResult smoke_simulation_ext_set_learned_preconditioner(SmokeSimulationContext context, void* network) {
    return smoke::guarded_call(context, [&](Context& c) {
        if (network) smoke::ensure_pcg_z(c);
        c.learned = static_cast<SmokeNnPreconditioner*>(network);
    });
}

SmokeSimulationResult smoke_simulation_ext_default_multilevel_options(SmokeSimulationMultilevelOptions* out) {
    if (!out) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    *out = smoke::default_multilevel_options();
    return SMOKE_SIMULATION_RESULT_OK;
}

SmokeSimulationResult smoke_simulation_ext_set_multilevel_options(SmokeSimulationContext context, const SmokeSimulationMultilevelOptions* o) {
    if (!o || !context) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    if (!smoke::multilevel_options_valid(*context, *o)) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    return smoke::guarded_call(context, [&](Context& c) {
        hip_check(hipStreamSynchronize(c.stream), "sync");   // buffers of the old hierarchy may be in use
        if (o->levels != c.mg.options.levels) smoke::multilevel_release(c);
        c.mg.options = *o;
        c.mg.built = 0;
    });
}

SmokeSimulationResult smoke_simulation_ext_get_multilevel_options(SmokeSimulationContext context, SmokeSimulationMultilevelOptions* out) {
    if (!context || !out) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    *out = context->mg.options;
    return SMOKE_SIMULATION_RESULT_OK;
}

SmokeSimulationResult smoke_simulation_ext_get_multilevel_info(SmokeSimulationContext context, SmokeSimulationMultilevelInfo* out) {
    if (!out) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    return smoke::guarded_call(context, [&](Context& c) {
        smoke::multilevel_build(c);
        SmokeSimulationMultilevelInfo info{};
        info.levels = c.mg.built;
        for (int l = 0; l < c.mg.built; ++l)
            for (int k = 0; k < 3; ++k) info.dims[l][k] = c.mg.level[l].n[k];
        info.captured = c.mg.captured;
        *out = info;
    });
}

SmokeSimulationResult smoke_simulation_ext_apply_multilevel(SmokeSimulationContext context, const float* r, float* z) {
    if (!r || !z || r == z) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    return smoke::guarded_call(context, [&](Context& c) { smoke::multilevel_apply(c, r, z); });
}

SmokeSimulationResult smoke_simulation_ext_export_level_flags(SmokeSimulationContext context, int32_t level, uint8_t* flags) {
    if (!flags) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    return smoke::guarded_call(context, [&](Context& c) {
        smoke::multilevel_build(c);
        if (level < 0 || level >= c.mg.built) throw smoke::HipError("no such level");
        const smoke::MgLevel& L = c.mg.level[level];
        hip_check(hipMemcpyAsync(flags, L.flags, size_t(L.cells), hipMemcpyDeviceToDevice, c.stream), "level flags");
    });
}

SmokeSimulationResult smoke_simulation_ext_time_multilevel(SmokeSimulationContext context, int32_t repetitions, SmokeSimulationMultilevelTiming* out) {
    if (!out || repetitions <= 0) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    return smoke::guarded_call(context, [&](Context& c) {
        smoke::ensure_pcg_z(c);
        *out = smoke::multilevel_timing(c, repetitions);
    });
}

SmokeSimulationResult smoke_simulation_ext_set_multilevel_capture(SmokeSimulationContext context, float* buffer, uint32_t capacity) {
    return smoke::guarded_call(context, [&](Context& c) {
        c.mg.capture = capacity > 0 ? buffer : nullptr;
        c.mg.capacity = buffer ? capacity : 0;
        c.mg.captured = 0;
    });
}

SmokeSimulationResult smoke_simulation_ext_measure_copy_bandwidth(uint64_t bytes, int32_t repetitions, double* out_gbps, double* out_ms) {
    if (!out_gbps || bytes < 16 || repetitions <= 0) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    try {
        *out_gbps = smoke::measure_copy(bytes, repetitions, out_ms);
        return SMOKE_SIMULATION_RESULT_OK;
    } catch (...) {
        return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    }
}

}  // extern "C"
