// HIP counterpart of the reference's timing CLI (benchmark.cu): the same
// flags, scenario and two output lines, driving the library only through the
// C ABI (visual-simulation-of-smoke-3d.h).  Extra flags select the extension
// solver (--tolerance, --jacobi, --adaptive-dt).
//
// Scenario (benchmark.cu:74-147): periodic x/z, no-slip floor, outflow top,
// an ellipsoidal emitter at (0.5, 0.12, 0.5) of the domain with radii
// (0.07, 0.05, 0.07) injecting 18 exp(-2.2 r^2) density and 36 exp(-2.2 r^2)
// temperature.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <string>
#include <vector>

#include "visual-simulation-of-smoke-3d-ext.h"

namespace {

bool ok_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    std::fprintf(stderr, "%s failed: %s\n", what, hipGetErrorString(e));
    return false;
}

bool ok_smoke(SmokeSimulationResult r, const char* what) {
    if (r == SMOKE_SIMULATION_RESULT_OK) return true;
    std::fprintf(stderr, "%s failed: %d\n", what, int(r));
    return false;
}

struct Args {
    int nx = 128, ny = 160, nz = 128, warmup = 24, steps = 128, pressure_iters = 64;
    float cell = 0.01f, dt = 1.0f / 90.0f, ambient = 0.0f, buoy_density = 0.15f, buoy_temperature = 1.2f, vorticity = 0.22f;
    bool cubic = true, tolerance = false, jacobi = false, adaptive = false;
    double tol = 1e-6;
};

int parse(int argc, char** argv, Args& a) {
    for (int i = 1; i < argc; ++i) {
        const std::string s = argv[i];
        auto val = [&](const char* flag) -> const char* {
            if (i + 1 < argc) return argv[++i];
            std::fprintf(stderr, "missing value for %s\n", flag);
            std::exit(EXIT_FAILURE);
        };
        if (s == "--nx") a.nx = std::atoi(val("--nx"));
        else if (s == "--ny") a.ny = std::atoi(val("--ny"));
        else if (s == "--nz") a.nz = std::atoi(val("--nz"));
        else if (s == "--warmup") a.warmup = std::atoi(val("--warmup"));
        else if (s == "--steps") a.steps = std::atoi(val("--steps"));
        else if (s == "--pressure-iters") a.pressure_iters = std::atoi(val("--pressure-iters"));
        else if (s == "--cell-size") a.cell = std::strtof(val("--cell-size"), nullptr);
        else if (s == "--dt") a.dt = std::strtof(val("--dt"), nullptr);
        else if (s == "--ambient-temperature") a.ambient = std::strtof(val("--ambient-temperature"), nullptr);
        else if (s == "--buoyancy-density") a.buoy_density = std::strtof(val("--buoyancy-density"), nullptr);
        else if (s == "--buoyancy-temperature") a.buoy_temperature = std::strtof(val("--buoyancy-temperature"), nullptr);
        else if (s == "--vorticity-confinement") a.vorticity = std::strtof(val("--vorticity-confinement"), nullptr);
        else if (s == "--linear-advection") a.cubic = false;
        else if (s == "--tolerance") a.tolerance = true, a.tol = std::strtod(val("--tolerance"), nullptr);
        else if (s == "--jacobi") a.jacobi = true;
        else if (s == "--adaptive-dt") a.adaptive = true;
        else if (s == "--help") {
            std::printf(
                "visual-simulation-of-smoke-benchmark [--nx N] [--ny N] [--nz N] [--warmup N] [--steps N]\n"
                "                        [--pressure-iters N] [--cell-size H] [--dt DT]\n"
                "                        [--ambient-temperature T0]\n"
                "                        [--buoyancy-density A] [--buoyancy-temperature B]\n"
                "                        [--vorticity-confinement E] [--linear-advection]\n"
                "                        [--tolerance TOL] [--jacobi] [--adaptive-dt]\n");
            return 1;
        } else {
            std::fprintf(stderr, "unknown argument: %s\n", argv[i]);
            return -1;
        }
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    Args a;
    const int pr = parse(argc, argv, a);
    if (pr != 0) return pr > 0 ? EXIT_SUCCESS : EXIT_FAILURE;

    SmokeSimulationConfig cfg{};
    cfg.nx = a.nx;
    cfg.ny = a.ny;
    cfg.nz = a.nz;
    cfg.cell_size = a.cell;
    cfg.dt = a.dt;
    cfg.pressure_iterations = a.pressure_iters;
    cfg.ambient_temperature = a.ambient;
    cfg.buoyancy_density_factor = a.buoy_density;
    cfg.buoyancy_temperature_factor = a.buoy_temperature;
    cfg.vorticity_confinement = a.vorticity;
    cfg.scalar_advection_mode = a.cubic ? SMOKE_SIMULATION_SCALAR_ADVECTION_MONOTONIC_CUBIC : SMOKE_SIMULATION_SCALAR_ADVECTION_LINEAR;
    cfg.flow_boundary.x_minus.type = cfg.flow_boundary.x_plus.type = SMOKE_SIMULATION_FLOW_BOUNDARY_PERIODIC;
    cfg.flow_boundary.z_minus.type = cfg.flow_boundary.z_plus.type = SMOKE_SIMULATION_FLOW_BOUNDARY_PERIODIC;
    cfg.flow_boundary.y_minus.type = SMOKE_SIMULATION_FLOW_BOUNDARY_NO_SLIP_WALL;
    cfg.flow_boundary.y_plus.type = SMOKE_SIMULATION_FLOW_BOUNDARY_OUTFLOW;
    for (SmokeSimulationScalarBoundaryConfig* s : {&cfg.density_boundary, &cfg.temperature_boundary}) {
        const float v = s == &cfg.density_boundary ? 0.0f : a.ambient;
        s->x_minus = s->x_plus = s->z_minus = s->z_plus = {SMOKE_SIMULATION_SCALAR_BOUNDARY_PERIODIC, v};
        s->y_minus = s->y_plus = {SMOKE_SIMULATION_SCALAR_BOUNDARY_FIXED_VALUE, v};
    }

    const size_t cells = size_t(a.nx) * size_t(a.ny) * size_t(a.nz);
    const size_t bytes = cells * sizeof(float);
    std::vector<float> rho_src(cells, 0.0f), t_src(cells, 0.0f);
    const float ex = a.nx * a.cell, ey = a.ny * a.cell, ez = a.nz * a.cell;
    const float cx = 0.5f * ex, cy = 0.12f * ey, cz = 0.5f * ez, rx = 0.07f * ex, ry = 0.05f * ey, rz = 0.07f * ez;
    for (int z = 0; z < a.nz; ++z)
        for (int y = 0; y < a.ny; ++y)
            for (int x = 0; x < a.nx; ++x) {
                const float dx = ((x + 0.5f) * a.cell - cx) / rx, dy = ((y + 0.5f) * a.cell - cy) / ry, dz = ((z + 0.5f) * a.cell - cz) / rz;
                const float r2 = dx * dx + dy * dy + dz * dz;
                if (r2 > 1.0f) continue;
                const size_t i = size_t(x) + size_t(a.nx) * (size_t(y) + size_t(a.ny) * size_t(z));
                rho_src[i] = 18.0f * std::exp(-2.2f * r2);
                t_src[i] = 36.0f * std::exp(-2.2f * r2);
            }

    hipStream_t stream = nullptr;
    if (!ok_hip(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreateWithFlags")) return EXIT_FAILURE;
    SmokeSimulationContextCreateDesc desc{};
    desc.config = cfg;
    desc.stream = stream;
    desc.initial_density = 0.0f;
    desc.initial_temperature = a.ambient;
    SmokeSimulationContext ctx = nullptr;
    if (!ok_smoke(smoke_simulation_create_context_cuda(&desc, &ctx), "smoke_simulation_create_context_cuda")) return EXIT_FAILURE;
    if (a.tolerance || a.adaptive) {
        SmokeSimulationSolverOptions o{};
        smoke_simulation_ext_get_solver_options(ctx, &o);
        if (a.tolerance) {
            o.pressure_mode = SMOKE_SIMULATION_PRESSURE_TOLERANCE;
            o.relative_tolerance = a.tol;
            o.preconditioner = a.jacobi ? SMOKE_SIMULATION_PRECONDITIONER_JACOBI : SMOKE_SIMULATION_PRECONDITIONER_NONE;
        }
        o.adaptive_dt = a.adaptive ? 1u : 0u;
        o.dt_max = a.dt;
        if (!ok_smoke(smoke_simulation_ext_set_solver_options(ctx, &o), "smoke_simulation_ext_set_solver_options")) return EXIT_FAILURE;
    }
    float *d_rho = nullptr, *d_t = nullptr;
    if (!ok_hip(hipMalloc(&d_rho, bytes), "hipMalloc density_source") || !ok_hip(hipMalloc(&d_t, bytes), "hipMalloc temperature_source")) return EXIT_FAILURE;
    if (!ok_hip(hipMemcpyAsync(d_rho, rho_src.data(), bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync density_source")) return EXIT_FAILURE;
    if (!ok_hip(hipMemcpyAsync(d_t, t_src.data(), bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync temperature_source")) return EXIT_FAILURE;
    if (!ok_smoke(smoke_simulation_update_density_source_cuda(ctx, d_rho), "smoke_simulation_update_density_source_cuda")) return EXIT_FAILURE;
    if (!ok_smoke(smoke_simulation_update_temperature_source_cuda(ctx, d_t), "smoke_simulation_update_temperature_source_cuda")) return EXIT_FAILURE;

    hipEvent_t begin = nullptr, end = nullptr;
    if (!ok_hip(hipEventCreate(&begin), "hipEventCreate") || !ok_hip(hipEventCreate(&end), "hipEventCreate")) return EXIT_FAILURE;
    for (int s = 0; s < a.warmup; ++s)
        if (!ok_smoke(smoke_simulation_step_cuda(ctx), "smoke_simulation_step_cuda warmup")) return EXIT_FAILURE;
    if (!ok_hip(hipStreamSynchronize(stream), "hipStreamSynchronize warmup")) return EXIT_FAILURE;
    float ms = 0.0f;
    if (!ok_hip(hipEventRecord(begin, stream), "hipEventRecord")) return EXIT_FAILURE;
    for (int s = 0; s < a.steps; ++s)
        if (!ok_smoke(smoke_simulation_step_cuda(ctx), "smoke_simulation_step_cuda")) return EXIT_FAILURE;
    if (!ok_hip(hipEventRecord(end, stream), "hipEventRecord") || !ok_hip(hipEventSynchronize(end), "hipEventSynchronize")) return EXIT_FAILURE;
    if (!ok_hip(hipEventElapsedTime(&ms, begin, end), "hipEventElapsedTime")) return EXIT_FAILURE;

    std::vector<float> density(cells, 0.0f);
    SmokeSimulationView view{};
    const SmokeSimulationViewRequest req{SMOKE_SIMULATION_VIEW_DENSITY, stream};
    if (!ok_smoke(smoke_simulation_get_view_cuda(ctx, &req, &view), "smoke_simulation_get_view_cuda")) return EXIT_FAILURE;
    if (!ok_hip(hipMemcpyAsync(density.data(), view.data0, bytes, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync density_export")) return EXIT_FAILURE;
    if (!ok_hip(hipStreamSynchronize(stream), "hipStreamSynchronize export")) return EXIT_FAILURE;

    const double avg = a.steps > 0 ? double(ms) / a.steps : 0.0;
    const double mlups = ms > 0.0f ? double(cells) * a.steps / (double(ms) * 1000.0) : 0.0;
    const float total = std::accumulate(density.begin(), density.end(), 0.0f);
    const float peak = density.empty() ? 0.0f : *std::max_element(density.begin(), density.end());
    std::printf(
        "benchmark grid=%dx%dx%d warmup=%d steps=%d dt=%.6f ambient=%.3f buoyancy_density=%.3f buoyancy_temperature=%.3f vorticity_confinement=%.3f "
        "advection=%s\n",
        a.nx, a.ny, a.nz, a.warmup, a.steps, a.dt, a.ambient, a.buoy_density, a.buoy_temperature, a.vorticity, a.cubic ? "monotonic_cubic" : "linear");
    std::printf("timing total_ms=%.3f avg_step_ms=%.6f mlups=%.3f total_density=%.6f peak_density=%.6f\n", double(ms), avg, mlups, total, peak);

    hipEventDestroy(end);
    hipEventDestroy(begin);
    hipFree(d_t);
    hipFree(d_rho);
    smoke_simulation_destroy_context_cuda(ctx);
    hipStreamDestroy(stream);
    return EXIT_SUCCESS;
}
