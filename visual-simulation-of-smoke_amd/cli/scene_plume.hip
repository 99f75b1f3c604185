// The reference visualizer's scene as a C-ABI caller (vulkan-app/
// scene_plume.cpp:180-392, scene_plume.ixx:34-91), without the window and
// renderer: it drives the library exactly as the scene does each frame.
//   * rebuild (scene_plume.cpp:223-318): a caller-owned non-blocking stream;
//     the context created on it with the initial temperature at ambient; an
//     ellipsoidal emitter at (0.5, 0.1, 0.5) of the domain with radii
//     (0.12, 0.08, 0.12), weight exp(-2.2 r^2), density source 18 w and
//     temperature source 34 w, uploaded from host arrays with async copies and
//     handed over with update_*; a zero user force.
//   * step (scene_plume.cpp:321-361): per step a force of speed 22 U(0.82,
//     1.2) along (U(-0.28, 0.28), 1, U(-0.28, 0.28)) normalised, times the
//     emitter weight in every cell, copied to device buffers and set with
//     update_force_cuda, then step_cuda.
//   * export (scene_plume.cpp:363-392): get_view with the caller's stream as
//     consumer, then device-to-device copies of the density, temperature and
//     cell-centred velocity views into the caller's buffers (as the Vulkan
//     app does into its imported images), and a device-to-host copy of the
//     velocity.
// The reference seeds its generator from std::random_device; here the seed is
// a flag so that a run can be repeated.  The run's inputs and exports go to a
// binary file for checking ("SMOKESC1": header, emitter weights, per-step
// force components, then the exported density, temperature and velocity
// x/y/z as FP32 cell arrays).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "visual-simulation-of-smoke-3d.h"

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
    int nx = 96, ny = 128, nz = 96, steps = 60;
    unsigned seed = 1;
    std::string out;
};

SmokeSimulationConfig scene_config(const Args& a) {
    // scene_plume.ixx:34-78
    SmokeSimulationConfig c{};
    c.nx = a.nx;
    c.ny = a.ny;
    c.nz = a.nz;
    c.cell_size = 0.01f;
    c.dt = 1.0f / 90.0f;
    c.pressure_iterations = 96;
    c.ambient_temperature = 0.0f;
    c.buoyancy_density_factor = 0.8f;
    c.buoyancy_temperature_factor = 0.75f;
    c.vorticity_confinement = 0.20f;
    c.scalar_advection_mode = SMOKE_SIMULATION_SCALAR_ADVECTION_MONOTONIC_CUBIC;
    const SmokeSimulationFlowBoundaryFaceDesc periodic{SMOKE_SIMULATION_FLOW_BOUNDARY_PERIODIC, 0.0f, 0.0f, 0.0f, 0.0f};
    c.flow_boundary.x_minus = c.flow_boundary.x_plus = c.flow_boundary.z_minus = c.flow_boundary.z_plus = periodic;
    c.flow_boundary.y_minus = {SMOKE_SIMULATION_FLOW_BOUNDARY_NO_SLIP_WALL, 0.0f, 0.0f, 0.0f, 0.0f};
    c.flow_boundary.y_plus = {SMOKE_SIMULATION_FLOW_BOUNDARY_OUTFLOW, 0.0f, 0.0f, 0.0f, 0.0f};
    for (SmokeSimulationScalarBoundaryConfig* s : {&c.density_boundary, &c.temperature_boundary}) {
        s->x_minus = s->x_plus = s->z_minus = s->z_plus = {SMOKE_SIMULATION_SCALAR_BOUNDARY_PERIODIC, 0.0f};
        s->y_minus = s->y_plus = {SMOKE_SIMULATION_SCALAR_BOUNDARY_FIXED_VALUE, 0.0f};
    }
    return c;
}

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
        else if (s == "--steps") a.steps = std::atoi(val("--steps"));
        else if (s == "--seed") a.seed = unsigned(std::strtoul(val("--seed"), nullptr, 10));
        else if (s == "--out") a.out = val("--out");
        else if (s == "--help") {
            std::printf("visual-simulation-of-smoke-scene [--nx N] [--ny N] [--nz N] [--steps N] [--seed S] [--out FILE]\n");
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
    if (a.nx <= 0 || a.ny <= 0 || a.nz <= 0 || a.steps < 0) {
        std::fprintf(stderr, "invalid grid or step count\n");
        return EXIT_FAILURE;
    }
    const SmokeSimulationConfig config = scene_config(a);
    hipStream_t stream = nullptr;
    if (!ok_hip(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreateWithFlags")) return EXIT_FAILURE;

    // rebuild: context on the caller's stream (scene_plume.cpp:240-245)
    SmokeSimulationContextCreateDesc desc{};
    desc.config = config;
    desc.stream = stream;
    desc.initial_density = 0.0f;
    desc.initial_temperature = config.ambient_temperature;
    SmokeSimulationContext ctx = nullptr;
    if (!ok_smoke(smoke_simulation_create_context_cuda(&desc, &ctx), "smoke_simulation_create_context_cuda")) return EXIT_FAILURE;

    const size_t cells = size_t(a.nx) * size_t(a.ny) * size_t(a.nz);
    const size_t bytes = cells * sizeof(float);
    const float h = config.cell_size;
    std::vector<float> weight(cells, 0.0f), rho_src(cells, 0.0f), t_src(cells, 0.0f), fx(cells, 0.0f), fy(cells, 0.0f), fz(cells, 0.0f);
    const float ex = float(a.nx) * h * 0.50f, ey = float(a.ny) * h * 0.10f, ez = float(a.nz) * h * 0.50f;
    const float rx = float(a.nx) * h * 0.12f, ry = float(a.ny) * h * 0.08f, rz = float(a.nz) * h * 0.12f;
    for (int z = 0; z < a.nz; ++z)
        for (int y = 0; y < a.ny; ++y)
            for (int x = 0; x < a.nx; ++x) {
                const size_t i = size_t(x) + size_t(a.nx) * (size_t(y) + size_t(a.ny) * size_t(z));
                const float dx = ((float(x) + 0.5f) * h - ex) / rx;
                const float dy = ((float(y) + 0.5f) * h - ey) / ry;
                const float dz = ((float(z) + 0.5f) * h - ez) / rz;
                const float r2 = dx * dx + dy * dy + dz * dz;
                if (r2 > 1.0f) continue;
                const float w = std::exp(-2.2f * r2);
                weight[i] = w;
                rho_src[i] = 18.0f * w;
                t_src[i] = 34.0f * w;
            }
    float *d_rho_src = nullptr, *d_t_src = nullptr, *d_f[3] = {}, *d_export[5] = {};
    bool ok = ok_hip(hipMalloc(&d_rho_src, bytes), "hipMalloc") && ok_hip(hipMalloc(&d_t_src, bytes), "hipMalloc");
    for (float*& p : d_f) ok = ok && ok_hip(hipMalloc(&p, bytes), "hipMalloc");
    for (float*& p : d_export) ok = ok && ok_hip(hipMalloc(&p, bytes), "hipMalloc");
    ok = ok && ok_hip(hipMemcpyAsync(d_rho_src, rho_src.data(), bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync") &&
         ok_hip(hipMemcpyAsync(d_t_src, t_src.data(), bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync") &&
         ok_hip(hipMemcpyAsync(d_f[0], fx.data(), bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync") &&
         ok_hip(hipMemcpyAsync(d_f[1], fy.data(), bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync") &&
         ok_hip(hipMemcpyAsync(d_f[2], fz.data(), bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync") &&
         ok_smoke(smoke_simulation_update_density_source_cuda(ctx, d_rho_src), "update_density_source") &&
         ok_smoke(smoke_simulation_update_temperature_source_cuda(ctx, d_t_src), "update_temperature_source") &&
         ok_smoke(smoke_simulation_update_force_cuda(ctx, d_f[0], d_f[1], d_f[2]), "update_force");

    // steps (scene_plume.cpp:321-361)
    std::mt19937 engine{a.seed};
    std::uniform_real_distribution<float> speed_distribution{0.82f, 1.20f};
    std::uniform_real_distribution<float> direction_distribution{-0.28f, 0.28f};
    std::vector<float> forces;
    double call_ms = 0.0;
    for (int s = 0; ok && s < a.steps; ++s) {
        const auto begin = std::chrono::steady_clock::now();
        const float lateral_x = direction_distribution(engine);
        const float lateral_z = direction_distribution(engine);
        const float speed = 22.0f * speed_distribution(engine);
        const float length = std::sqrt(lateral_x * lateral_x + 1.0f + lateral_z * lateral_z);
        const float f[3] = {speed * (lateral_x / length), speed * (1.0f / length), speed * (lateral_z / length)};
        forces.insert(forces.end(), f, f + 3);
        for (size_t i = 0; i < cells; ++i) {
            fx[i] = weight[i] * f[0];
            fy[i] = weight[i] * f[1];
            fz[i] = weight[i] * f[2];
        }
        ok = ok_hip(hipMemcpyAsync(d_f[0], fx.data(), bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync") &&
             ok_hip(hipMemcpyAsync(d_f[1], fy.data(), bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync") &&
             ok_hip(hipMemcpyAsync(d_f[2], fz.data(), bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync") &&
             ok_smoke(smoke_simulation_update_force_cuda(ctx, d_f[0], d_f[1], d_f[2]), "update_force") &&
             ok_smoke(smoke_simulation_step_cuda(ctx), "step");
        // the host arrays are rewritten next step: the copies out of them
        // must have run (the scene relies on the same ordering through its
        // frame fences)
        ok = ok && ok_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
        call_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - begin).count();
    }

    // export (scene_plume.cpp:363-392)
    auto export_view = [&](uint32_t kind, float* const* dst, int components) {
        SmokeSimulationView view{};
        const SmokeSimulationViewRequest request{kind, stream};
        if (!ok_smoke(smoke_simulation_get_view_cuda(ctx, &request, &view), "get_view")) return false;
        const void* src[3] = {view.data0, view.data1, view.data2};
        for (int c = 0; c < components; ++c)
            if (!ok_hip(hipMemcpyAsync(dst[c], src[c], bytes, hipMemcpyDeviceToDevice, stream), "export copy")) return false;
        return true;
    };
    std::vector<float> host_velocity(3 * cells, 0.0f);
    ok = ok && export_view(SMOKE_SIMULATION_VIEW_DENSITY, &d_export[0], 1) && export_view(SMOKE_SIMULATION_VIEW_TEMPERATURE, &d_export[1], 1) &&
         export_view(SMOKE_SIMULATION_VIEW_FLOW_VELOCITY, &d_export[2], 3);
    if (ok) {
        SmokeSimulationView view{};
        const SmokeSimulationViewRequest request{SMOKE_SIMULATION_VIEW_FLOW_VELOCITY, stream};
        ok = ok_smoke(smoke_simulation_get_view_cuda(ctx, &request, &view), "get_view");
        const void* src[3] = {view.data0, view.data1, view.data2};
        for (int c = 0; ok && c < 3; ++c)
            ok = ok_hip(hipMemcpyAsync(host_velocity.data() + c * cells, src[c], bytes, hipMemcpyDeviceToHost, stream), "velocity host copy");
    }
    std::vector<float> exported(5 * cells, 0.0f);
    for (int c = 0; ok && c < 5; ++c)
        ok = ok_hip(hipMemcpyAsync(exported.data() + c * cells, d_export[c], bytes, hipMemcpyDeviceToHost, stream), "readback");
    ok = ok && ok_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
    // the host copy of the velocity equals the device export
    for (size_t i = 0; ok && i < 3 * cells; ++i)
        if (host_velocity[i] != exported[2 * cells + i]) {
            std::fprintf(stderr, "host and device velocity exports differ at %zu\n", i);
            ok = false;
        }
    if (ok && !a.out.empty()) {
        FILE* f = std::fopen(a.out.c_str(), "wb");
        const int32_t head[5] = {a.nx, a.ny, a.nz, a.steps, int32_t(a.seed)};
        ok = f && std::fwrite("SMOKESC1", 1, 8, f) == 8 && std::fwrite(head, 4, 5, f) == 5 && std::fwrite(weight.data(), 4, cells, f) == cells &&
             std::fwrite(forces.data(), 4, forces.size(), f) == forces.size() && std::fwrite(exported.data(), 4, exported.size(), f) == exported.size();
        if (f) ok = std::fclose(f) == 0 && ok;
        if (!ok) std::fprintf(stderr, "cannot write %s\n", a.out.c_str());
    }
    if (ok) std::printf("scene_plume: %d steps on %dx%dx%d, %.3f ms per step call\n", a.steps, a.nx, a.ny, a.nz, a.steps ? call_ms / a.steps : 0.0);
    smoke_simulation_destroy_context_cuda(ctx);
    for (float* p : {d_rho_src, d_t_src, d_f[0], d_f[1], d_f[2], d_export[0], d_export[1], d_export[2], d_export[3], d_export[4]})
        if (p) (void)hipFree(p);
    (void)hipStreamDestroy(stream);
    return ok ? EXIT_SUCCESS : EXIT_FAILURE;
}
