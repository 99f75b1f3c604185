"""CPU-only checks of the C-ABI boundary: the built library exports every
symbol include/*.h declares, the ctypes mirror matches the C layouts (gcc
compiles the header), and the error paths of backend-cuda.cu:1833-1863,
2074-2077, 2161-2164 and 2174-2177, 2281 behave the same without a GPU."""
import ctypes as C
import subprocess
import textwrap
from pathlib import Path

import pytest

import smoke_abi as abi

REPO = Path(__file__).resolve().parents[1]


def _lib():
    if not abi.LIB_PATH.exists():
        pytest.skip("library not built (run __graft_entry__.build())")
    return abi.load_library()


def test_library_exports_every_header_symbol():
    _lib()
    names = abi.abi_symbols_from_headers(REPO / "include")
    assert len(names) >= 12
    out = subprocess.run(["nm", "-D", "--defined-only", str(abi.LIB_PATH)], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    # the 12 reference entry points (visual-simulation-of-smoke-3d.h:103-114, 156)
    ref = ["smoke_simulation_create_context_cuda", "smoke_simulation_destroy_context_cuda", "smoke_simulation_update_density_cuda",
           "smoke_simulation_update_density_source_cuda", "smoke_simulation_update_temperature_cuda",
           "smoke_simulation_update_temperature_source_cuda", "smoke_simulation_update_force_cuda",
           "smoke_simulation_update_occupancy_cuda", "smoke_simulation_update_solid_velocity_cuda",
           "smoke_simulation_update_solid_temperature_cuda", "smoke_simulation_step_cuda", "smoke_simulation_get_view_cuda"]
    assert all(r in exported for r in ref)


def test_struct_layouts_match_c_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(textwrap.dedent("""
        #include <stdio.h>
        #include <stddef.h>
        #include "visual-simulation-of-smoke-3d-ext.h"
        #define S(t) printf(#t " %zu\\n", sizeof(t));
        #define O(t, f) printf(#t "." #f " %zu\\n", offsetof(t, f));
        int main(void) {
            S(SmokeSimulationConfig) S(SmokeSimulationContextCreateDesc) S(SmokeSimulationView)
            S(SmokeSimulationViewRequest) S(SmokeSimulationFlowBoundaryFaceDesc) S(SmokeSimulationScalarBoundaryFaceDesc)
            S(SmokeSimulationSolverOptions) S(SmokeSimulationSolverStats) S(SmokeSimulationHostState)
            S(SmokeSimulationKernelTiming) S(SmokeSimulationPlumeSource)
            O(SmokeSimulationConfig, scalar_advection_mode) O(SmokeSimulationConfig, flow_boundary)
            O(SmokeSimulationConfig, density_boundary) O(SmokeSimulationConfig, temperature_boundary)
            O(SmokeSimulationContextCreateDesc, stream) O(SmokeSimulationContextCreateDesc, initial_temperature)
            O(SmokeSimulationView, row_stride_bytes) O(SmokeSimulationView, data0)
            O(SmokeSimulationSolverOptions, relative_tolerance) O(SmokeSimulationSolverStats, relative_residual)
            O(SmokeSimulationSolverStats, steps)
            return 0;
        }"""))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{REPO / 'include'}", str(src), "-o", str(exe)], check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines())
    got = {k: int(v) for k, v in got.items()}
    # sizes the reference header gives on LP64 (measured against visual-simulation-of-smoke-3d.h)
    assert got["SmokeSimulationConfig"] == 260
    assert got["SmokeSimulationContextCreateDesc"] == 280
    assert got["SmokeSimulationView"] == 56
    assert got["SmokeSimulationConfig.scalar_advection_mode"] == 40
    assert got["SmokeSimulationConfig.flow_boundary"] == 44
    assert got["SmokeSimulationConfig.density_boundary"] == 164
    assert got["SmokeSimulationConfig.temperature_boundary"] == 212
    assert got["SmokeSimulationContextCreateDesc.stream"] == 264
    assert got["SmokeSimulationContextCreateDesc.initial_temperature"] == 276
    assert got["SmokeSimulationView.row_stride_bytes"] == 16
    assert got["SmokeSimulationView.data0"] == 32
    mirror = {"SmokeSimulationConfig": abi.Config, "SmokeSimulationContextCreateDesc": abi.CreateDesc, "SmokeSimulationView": abi.View,
              "SmokeSimulationViewRequest": abi.ViewRequest, "SmokeSimulationFlowBoundaryFaceDesc": abi.FlowFace,
              "SmokeSimulationScalarBoundaryFaceDesc": abi.ScalarFace, "SmokeSimulationSolverOptions": abi.SolverOptions,
              "SmokeSimulationSolverStats": abi.SolverStats, "SmokeSimulationHostState": abi.HostState,
              "SmokeSimulationKernelTiming": abi.KernelTiming, "SmokeSimulationPlumeSource": abi.PlumeSource}
    for name, cls in mirror.items():
        assert C.sizeof(cls) == got[name], name
    assert abi.SolverStats.relative_residual.offset == got["SmokeSimulationSolverStats.relative_residual"]
    assert abi.SolverStats.steps.offset == got["SmokeSimulationSolverStats.steps"]


def _desc(**kw):
    d = abi.CreateDesc()
    d.config = abi.make_config(8, 8, 8, **kw)
    return d


def test_create_rejects_invalid_configs_before_touching_the_device():
    L = _lib()
    out = C.c_void_p()
    assert L.smoke_simulation_create_context_cuda(None, C.byref(out)) == abi.RESULT_BACKEND_FAILURE
    assert L.smoke_simulation_create_context_cuda(C.byref(_desc()), None) == abi.RESULT_BACKEND_FAILURE
    bad = []
    d = _desc(); d.config.nx = 0; bad.append(d)
    d = _desc(); d.config.cell_size = 0.0; bad.append(d)
    d = _desc(); d.config.dt = -1.0; bad.append(d)
    d = _desc(); d.config.pressure_iterations = 0; bad.append(d)
    d = _desc(); d.config.flow_boundary.x_minus.type = abi.PERIODIC; bad.append(d)        # unpaired periodic face
    d = _desc(); d.config.density_boundary.z_plus.type = abi.SCALAR_PERIODIC; bad.append(d)
    d = _desc(); d.config.flow_boundary.y_plus.type = 7; bad.append(d)                     # unknown face type
    d = _desc(); d.config.temperature_boundary.x_minus.type = 9; bad.append(d)
    for d in bad:
        assert L.smoke_simulation_create_context_cuda(C.byref(d), C.byref(out)) == abi.RESULT_BACKEND_FAILURE


def test_null_handles():
    L = _lib()
    assert L.smoke_simulation_destroy_context_cuda(None) == abi.RESULT_OK
    assert L.smoke_simulation_step_cuda(None) == abi.RESULT_BACKEND_FAILURE
    req, view = abi.ViewRequest(0, None), abi.View()
    assert L.smoke_simulation_get_view_cuda(None, C.byref(req), C.byref(view)) == abi.RESULT_BACKEND_FAILURE
    assert L.smoke_simulation_ext_run_stages(None, 0, 4) == abi.RESULT_BACKEND_FAILURE
    opts = abi.SolverOptions()
    assert L.smoke_simulation_ext_default_solver_options(C.byref(opts)) == abi.RESULT_OK
    assert opts.pressure_mode == abi.PRESSURE_FIXED_ITERATIONS and opts.relative_tolerance == 1e-6


def test_create_without_a_gpu_reports_backend_failure():
    L = _lib()
    hip = abi.Hip()
    if hip.device_count() > 0:
        pytest.skip("a GPU is present")
    out = C.c_void_p()
    assert L.smoke_simulation_create_context_cuda(C.byref(_desc()), C.byref(out)) == abi.RESULT_BACKEND_FAILURE
    assert not out.value
