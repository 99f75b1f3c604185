"""ctypes mirror of the smoke-step C ABI (include/visual-simulation-of-smoke-3d.h)
and of its extensions (include/visual-simulation-of-smoke-3d-ext.h).

This is host plumbing for tests and the benchmark: it loads the HIP library
built in-tree (lib/libvisual-simulation-of-smoke.so) and refuses to run
without it -- there is no CPU fallback.  Device buffers are allocated through
the HIP runtime (libamdhip64) so that the update_* entry points receive device
pointers exactly as the reference's callers pass them (benchmark.cu:163-172).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "lib" / "libvisual-simulation-of-smoke.so"

# ---------------------------------------------------------------- enums
RESULT_OK, RESULT_OUT_OF_MEMORY, RESULT_BACKEND_FAILURE = 0, 1, 2
NO_SLIP_WALL, FREE_SLIP_WALL, OUTFLOW, PERIODIC = 0, 1, 2, 3
FIXED_VALUE, ZERO_FLUX, SCALAR_PERIODIC = 0, 1, 2
ADVECTION_LINEAR, ADVECTION_MONOTONIC_CUBIC = 0, 1

VIEW_KINDS = {
    "DENSITY": 0, "DENSITY_SOURCE": 1, "TEMPERATURE": 2, "TEMPERATURE_SOURCE": 3, "FORCE": 4,
    "SOLID_VELOCITY": 5, "SOLID_TEMPERATURE": 6, "FLOW_VELOCITY": 7, "FLOW_VELOCITY_MAGNITUDE": 8,
    "FLOW_PRESSURE": 9, "FLOW_PRESSURE_RHS": 10, "FLOW_DIVERGENCE": 11, "FLOW_VORTICITY": 12,
    "FLOW_VORTICITY_MAGNITUDE": 13, "OCCUPANCY": 14,
}
LAYOUT_F32_3D, LAYOUT_F32_3D_SOA3 = 0, 1

PRESSURE_FIXED_ITERATIONS, PRESSURE_TOLERANCE = 0, 1
PRECONDITIONER_NONE, PRECONDITIONER_JACOBI, PRECONDITIONER_LEARNED, PRECONDITIONER_MULTILEVEL = 0, 1, 2, 3
COARSE_NETWORK, COARSE_JACOBI = 0, 1
NN_LOSS_RESIDUAL, NN_LOSS_ENERGY = 0, 1
STAGE_FORCES, STAGE_ADVECT_VELOCITY, STAGE_PRESSURE, STAGE_PROJECT, STAGE_SCALARS = range(5)


# ---------------------------------------------------------------- structs
class FlowFace(C.Structure):
    _fields_ = [("type", C.c_uint32), ("velocity_x", C.c_float), ("velocity_y", C.c_float),
                ("velocity_z", C.c_float), ("pressure", C.c_float)]


class FlowBoundary(C.Structure):
    _fields_ = [(n, FlowFace) for n in ("x_minus", "x_plus", "y_minus", "y_plus", "z_minus", "z_plus")]


class ScalarFace(C.Structure):
    _fields_ = [("type", C.c_uint32), ("value", C.c_float)]


class ScalarBoundary(C.Structure):
    _fields_ = [(n, ScalarFace) for n in ("x_minus", "x_plus", "y_minus", "y_plus", "z_minus", "z_plus")]


class Config(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("cell_size", C.c_float), ("dt", C.c_float), ("pressure_iterations", C.c_int32),
        ("ambient_temperature", C.c_float), ("buoyancy_density_factor", C.c_float),
        ("buoyancy_temperature_factor", C.c_float), ("vorticity_confinement", C.c_float),
        ("scalar_advection_mode", C.c_uint32),
        ("flow_boundary", FlowBoundary), ("density_boundary", ScalarBoundary),
        ("temperature_boundary", ScalarBoundary),
    ]


class CreateDesc(C.Structure):
    _fields_ = [("config", Config), ("stream", C.c_void_p), ("initial_density", C.c_float),
                ("initial_temperature", C.c_float)]


class ViewRequest(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("consumer_stream", C.c_void_p)]


class View(C.Structure):
    _fields_ = [("layout", C.c_uint32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("row_stride_bytes", C.c_uint64), ("slice_stride_bytes", C.c_uint64),
                ("data0", C.c_void_p), ("data1", C.c_void_p), ("data2", C.c_void_p)]


class SolverOptions(C.Structure):
    _fields_ = [("pressure_mode", C.c_uint32), ("preconditioner", C.c_uint32),
                ("relative_tolerance", C.c_double), ("max_iterations", C.c_int32),
                ("max_refinements", C.c_int32), ("warm_start", C.c_uint32), ("adaptive_dt", C.c_uint32),
                ("cfl", C.c_float), ("dt_max", C.c_float), ("check_interval", C.c_int32),
                ("reserved", C.c_int32)]


class SolverStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("refinements", C.c_int32), ("converged", C.c_int32),
                ("anchor", C.c_int32), ("relative_residual", C.c_double), ("rhs_norm", C.c_double),
                ("dt", C.c_float), ("max_face_speed", C.c_float), ("steps", C.c_uint64),
                ("learned_iterations", C.c_int32), ("learned_fallback", C.c_int32)]


class HostState(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "velocity_x", "velocity_y", "velocity_z", "density", "temperature", "pressure",
        "velocity_star_x", "velocity_star_y", "velocity_star_z")]


class KernelTiming(C.Structure):
    _fields_ = [("pcg_iterations", C.c_uint64), ("pcg_timed_iterations", C.c_uint64),
                ("pcg_direction_ms", C.c_double), ("pcg_curvature_ms", C.c_double), ("pcg_update_ms", C.c_double),
                ("pcg_reduction_ms", C.c_double),
                ("step_ms", C.c_double), ("steps", C.c_uint64), ("mg_sweep_ms", C.c_double), ("mg_sweeps", C.c_uint64)]


class LaunchShape(C.Structure):
    _fields_ = [("threads_x", C.c_int32), ("threads_y", C.c_int32), ("threads_z", C.c_int32), ("cells_per_thread", C.c_int32)]

    def tuple(self):
        return (self.threads_x, self.threads_y, self.threads_z, self.cells_per_thread)


class LaunchReport(C.Structure):
    _fields_ = [("pcg_curvature", LaunchShape), ("pcg_update", LaunchShape), ("advect", LaunchShape * 4),
                ("pcg_curvature_us", C.c_float * 2), ("pcg_update_us", C.c_float * 2), ("advect_us", (C.c_float * 2) * 4),
                ("candidates_timed", C.c_int32), ("from_cache", C.c_int32)]


class TrainingCrops(C.Structure):
    _fields_ = [("path", C.c_char_p), ("origins", C.c_void_p), ("count", C.c_uint32), ("size", C.c_uint32 * 3), ("step", C.c_uint32)]


class DeviceCrops(C.Structure):
    _fields_ = [("origins", C.c_void_p), ("count", C.c_uint32), ("size", C.c_uint32 * 3), ("r", C.c_void_p), ("flags", C.c_void_p)]


class MultilevelOptions(C.Structure):
    _fields_ = [("levels", C.c_int32), ("pre_sweeps", C.c_int32), ("post_sweeps", C.c_int32), ("omega", C.c_float),
                ("coarse_solver", C.c_uint32), ("coarse_sweeps", C.c_int32), ("reserved", C.c_int32 * 2)]


class MultilevelInfo(C.Structure):
    _fields_ = [("levels", C.c_int32), ("dims", (C.c_int32 * 3) * 8), ("captured", C.c_uint32)]


class MultilevelTiming(C.Structure):
    _fields_ = [("sweep_us", C.c_float), ("restrict_us", C.c_float), ("scale_us", C.c_float), ("prolong_us", C.c_float),
                ("vcycle_us", C.c_float), ("vector_rows", C.c_int32), ("repetitions", C.c_int32), ("reserved", C.c_int32)]


class PlumeSource(C.Structure):
    _fields_ = [("center_x", C.c_float), ("center_y", C.c_float), ("center_z", C.c_float),
                ("radius", C.c_float), ("density_rate", C.c_float), ("temperature_rate", C.c_float),
                ("temperature_noise", C.c_float), ("seed", C.c_uint32)]


class CheckpointInfo(C.Structure):
    _fields_ = [("version", C.c_uint32), ("dims", C.c_int32 * 3), ("field_count", C.c_uint32), ("solves_done", C.c_int32),
                ("error", C.c_char * 96)]


class BatchInfo(C.Structure):
    _fields_ = [("fused", C.c_int32), ("graph", C.c_int32), ("steps", C.c_uint64), ("launches_per_step", C.c_uint64),
                ("launches_per_iteration", C.c_uint64)]


class BatchCreateDesc(C.Structure):
    _fields_ = [("plumes", C.c_void_p), ("count", C.c_uint32), ("lanes", C.c_int32)]


# ---------------------------------------------------------------- config helpers
def make_config(nx, ny, nz, *, cell_size=None, dt=0.01, pressure_iterations=64, ambient=0.0,
                buoyancy_density=0.05, buoyancy_temperature=0.25, vorticity=2.0, cubic=True,
                flow=None, density_bc=None, temperature_bc=None) -> Config:
    """Builds a Config; `flow` is six (type, vx, vy, vz, p) tuples (x-,x+,y-,y+,z-,z+),
    `density_bc` / `temperature_bc` six (type, value) tuples.  Defaults: closed
    no-slip box, FIXED_VALUE scalars at 0 (density) / ambient (temperature)."""
    cfg = Config()
    cfg.nx, cfg.ny, cfg.nz = nx, ny, nz
    cfg.cell_size = cell_size if cell_size is not None else 1.0 / nx
    cfg.dt = dt
    cfg.pressure_iterations = pressure_iterations
    cfg.ambient_temperature = ambient
    cfg.buoyancy_density_factor = buoyancy_density
    cfg.buoyancy_temperature_factor = buoyancy_temperature
    cfg.vorticity_confinement = vorticity
    cfg.scalar_advection_mode = ADVECTION_MONOTONIC_CUBIC if cubic else ADVECTION_LINEAR
    names = ("x_minus", "x_plus", "y_minus", "y_plus", "z_minus", "z_plus")
    flow = flow or [(NO_SLIP_WALL, 0.0, 0.0, 0.0, 0.0)] * 6
    for n, f in zip(names, flow):
        face = getattr(cfg.flow_boundary, n)
        face.type, face.velocity_x, face.velocity_y, face.velocity_z, face.pressure = f
    density_bc = density_bc or [(FIXED_VALUE, 0.0)] * 6
    temperature_bc = temperature_bc or [(FIXED_VALUE, ambient)] * 6
    for n, d, t in zip(names, density_bc, temperature_bc):
        getattr(cfg.density_boundary, n).type, getattr(cfg.density_boundary, n).value = d
        getattr(cfg.temperature_boundary, n).type, getattr(cfg.temperature_boundary, n).value = t
    return cfg


def face_shape(cfg: Config, comp: int):
    """numpy (z, y, x) shape of the MAC face array of component `comp`."""
    n = [cfg.nx, cfg.ny, cfg.nz]
    n[comp] += 1
    return (n[2], n[1], n[0])


def cell_shape(cfg: Config):
    return (cfg.nz, cfg.ny, cfg.nx)


def _ptr(a: np.ndarray | None):
    return None if a is None else C.c_void_p(a.ctypes.data)


# ---------------------------------------------------------------- HIP runtime plumbing
class Hip:
    """The handful of HIP runtime calls the harness needs (device buffers, streams, sync)."""

    H2D, D2H, D2D = 1, 2, 3

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        L = self.lib
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        L.hipDeviceSynchronize.argtypes = []
        L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        L.hipStreamDestroy.argtypes = [C.c_void_p]
        L.hipStreamSynchronize.argtypes = [C.c_void_p]
        L.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipMemset", "hipDeviceSynchronize", "hipStreamCreate",
                  "hipStreamDestroy", "hipStreamSynchronize", "hipGetDeviceCount"):
            getattr(L, f).restype = C.c_int

    def check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with hipError {rc}")

    def device_count(self) -> int:
        n = C.c_int(0)
        rc = self.lib.hipGetDeviceCount(C.byref(n))
        return n.value if rc == 0 else 0

    def malloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        self.check(self.lib.hipMalloc(C.byref(p), max(int(nbytes), 1)), "hipMalloc")
        return p.value

    def free(self, p: int):
        if p:
            self.check(self.lib.hipFree(C.c_void_p(p)), "hipFree")

    def upload(self, arr: np.ndarray) -> int:
        arr = np.ascontiguousarray(arr)
        p = self.malloc(arr.nbytes)
        self.check(self.lib.hipMemcpy(C.c_void_p(p), C.c_void_p(arr.ctypes.data), arr.nbytes, self.H2D), "hipMemcpy H2D")
        return p

    def copy(self, dst: int, src: int, nbytes: int):
        """Device-to-device copy (ordered on the null stream)."""
        self.check(self.lib.hipMemcpy(C.c_void_p(dst), C.c_void_p(src), int(nbytes), self.D2D), "hipMemcpy D2D")

    def download(self, p: int, shape, dtype=np.float32) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        self.check(self.lib.hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(p), out.nbytes, self.D2H), "hipMemcpy D2H")
        return out

    def synchronize(self):
        self.check(self.lib.hipDeviceSynchronize(), "hipDeviceSynchronize")


# ---------------------------------------------------------------- library
_LIB = None


def load_library(path: Path | str | None = None):
    """Loads the in-tree HIP build; raises if it has not been built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: build it with `python __graft_entry__.py build` (no CPU fallback exists)")
    L = C.CDLL(str(p))
    R = C.c_int
    sigs = {
        "smoke_simulation_create_context_cuda": [C.POINTER(CreateDesc), C.POINTER(C.c_void_p)],
        "smoke_simulation_destroy_context_cuda": [C.c_void_p],
        "smoke_simulation_update_density_cuda": [C.c_void_p, C.c_void_p],
        "smoke_simulation_update_density_source_cuda": [C.c_void_p, C.c_void_p],
        "smoke_simulation_update_temperature_cuda": [C.c_void_p, C.c_void_p],
        "smoke_simulation_update_temperature_source_cuda": [C.c_void_p, C.c_void_p],
        "smoke_simulation_update_force_cuda": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
        "smoke_simulation_update_occupancy_cuda": [C.c_void_p, C.c_void_p],
        "smoke_simulation_update_solid_velocity_cuda": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
        "smoke_simulation_update_solid_temperature_cuda": [C.c_void_p, C.c_void_p],
        "smoke_simulation_step_cuda": [C.c_void_p],
        "smoke_simulation_get_view_cuda": [C.c_void_p, C.POINTER(ViewRequest), C.POINTER(View)],
        "smoke_simulation_ext_default_solver_options": [C.POINTER(SolverOptions)],
        "smoke_simulation_ext_set_solver_options": [C.c_void_p, C.POINTER(SolverOptions)],
        "smoke_simulation_ext_get_solver_options": [C.c_void_p, C.POINTER(SolverOptions)],
        "smoke_simulation_ext_get_solver_stats": [C.c_void_p, C.POINTER(SolverStats)],
        "smoke_simulation_ext_import_state": [C.c_void_p, C.POINTER(HostState)],
        "smoke_simulation_ext_export_state": [C.c_void_p, C.POINTER(HostState)],
        "smoke_simulation_ext_export_pressure_f64": [C.c_void_p, C.c_void_p],
        "smoke_simulation_ext_run_stages": [C.c_void_p, C.c_int32, C.c_int32],
        "smoke_simulation_ext_save_checkpoint": [C.c_void_p, C.c_char_p],
        "smoke_simulation_ext_load_checkpoint": [C.c_void_p, C.c_char_p],
        "smoke_simulation_ext_inspect_checkpoint": [C.c_char_p, C.POINTER(CheckpointInfo)],
        "smoke_simulation_ext_set_kernel_timing": [C.c_void_p, C.c_int32],
        "smoke_simulation_ext_get_kernel_timing": [C.c_void_p, C.POINTER(KernelTiming)],
        "smoke_simulation_ext_measure_copy_bandwidth": [C.c_uint64, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)],
        "smoke_simulation_ext_plume_sources": [C.c_void_p, C.POINTER(PlumeSource), C.c_uint32],
        "smoke_simulation_ext_autotune_launches": [C.c_void_p, C.POINTER(LaunchReport)],
        "smoke_simulation_ext_reset_launches": [C.c_void_p],
        "smoke_simulation_ext_get_launches": [C.c_void_p, C.POINTER(LaunchReport)],
        "smoke_simulation_ext_set_launches": [C.c_void_p, C.POINTER(LaunchReport)],
        "smoke_simulation_ext_set_force_pass": [C.c_void_p, C.c_int32],
        "smoke_simulation_ext_write_training_crops": [C.c_void_p, C.POINTER(TrainingCrops)],
        "smoke_simulation_ext_copy_training_crops": [C.c_void_p, C.POINTER(DeviceCrops)],
        "smoke_simulation_ext_set_learned_preconditioner": [C.c_void_p, C.c_void_p],
        "smoke_simulation_ext_default_multilevel_options": [C.POINTER(MultilevelOptions)],
        "smoke_simulation_ext_set_multilevel_options": [C.c_void_p, C.POINTER(MultilevelOptions)],
        "smoke_simulation_ext_get_multilevel_options": [C.c_void_p, C.POINTER(MultilevelOptions)],
        "smoke_simulation_ext_get_multilevel_info": [C.c_void_p, C.POINTER(MultilevelInfo)],
        "smoke_simulation_ext_apply_multilevel": [C.c_void_p, C.c_void_p, C.c_void_p],
        "smoke_simulation_ext_export_level_flags": [C.c_void_p, C.c_int32, C.c_void_p],
        "smoke_simulation_ext_set_multilevel_capture": [C.c_void_p, C.c_void_p, C.c_uint32],
        "smoke_simulation_ext_time_multilevel": [C.c_void_p, C.c_int32, C.POINTER(MultilevelTiming)],
        "smoke_simulation_ext_create_batch": [C.POINTER(BatchCreateDesc), C.POINTER(C.c_void_p)],
        "smoke_simulation_ext_destroy_batch": [C.c_void_p],
        "smoke_simulation_ext_batch_context": [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p), C.POINTER(C.c_int32)],
        "smoke_simulation_ext_step_batch": [C.c_void_p, C.c_void_p, C.c_uint32],
        "smoke_simulation_ext_synchronize_batch": [C.c_void_p],
        "smoke_simulation_ext_get_batch_info": [C.c_void_p, C.POINTER(BatchInfo)],
        "smoke_simulation_ext_set_batch_fusion": [C.c_void_p, C.c_int32],
        "smoke_simulation_ext_batch_lanes": [C.POINTER(C.c_uint64), C.c_uint32, C.c_int32, C.POINTER(C.c_int32)],
    }
    for name, args in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = R
    if path is None:
        _LIB = L
    return L


def abi_symbols_from_headers(include_dir: Path) -> list[str]:
    """Every function name declared with SMOKE_SIMULATION_API in include/*.h."""
    import re
    names = []
    for h in sorted(Path(include_dir).glob("*.h")):
        names += re.findall(r"SMOKE_SIMULATION_API\s+\w+\s+(\w+)\s*\(", h.read_text())
    return names


class SmokeError(RuntimeError):
    pass


def check(rc: int, what: str):
    if rc != RESULT_OK:
        raise SmokeError(f"{what} returned {rc}")


class SmokeContext:
    """One simulation context (owns its stream) driven through the C ABI."""

    def __init__(self, cfg: Config, initial_density=0.0, initial_temperature=None, hip: Hip | None = None):
        self.lib = load_library()
        self.hip = hip or Hip()
        self.cfg = cfg
        desc = CreateDesc()
        desc.config = cfg
        desc.stream = None
        desc.initial_density = initial_density
        desc.initial_temperature = cfg.ambient_temperature if initial_temperature is None else initial_temperature
        h = C.c_void_p()
        check(self.lib.smoke_simulation_create_context_cuda(C.byref(desc), C.byref(h)), "create_context")
        self.handle = h
        self._owned = []

    def close(self):
        if self.handle:
            check(self.lib.smoke_simulation_destroy_context_cuda(self.handle), "destroy_context")
            self.handle = None
        for p in self._owned:
            self.hip.free(p)
        self._owned = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- device uploads through update_*
    def _dev(self, arr, dtype=np.float32):
        if arr is None:
            return None
        p = self.hip.upload(np.ascontiguousarray(arr, dtype=dtype))
        self._owned.append(p)
        return C.c_void_p(p)

    def _release(self):
        self.hip.synchronize()
        for p in self._owned:
            self.hip.free(p)
        self._owned = []

    def update(self, name: str, *arrays, dtype=np.float32):
        fn = getattr(self.lib, f"smoke_simulation_update_{name}_cuda")
        ptrs = [self._dev(a, dtype) for a in arrays]
        check(fn(self.handle, *ptrs), f"update_{name}")
        self._release()

    def set_occupancy(self, occ: np.ndarray | None):
        self.update("occupancy", occ, dtype=np.uint8)

    def step(self, count: int = 1):
        for _ in range(count):
            check(self.lib.smoke_simulation_step_cuda(self.handle), "step")

    def run_stages(self, first: int, last: int):
        check(self.lib.smoke_simulation_ext_run_stages(self.handle, first, last), "run_stages")

    def synchronize(self):
        self.hip.synchronize()

    def view(self, kind: str):
        req = ViewRequest(VIEW_KINDS[kind], None)
        v = View()
        check(self.lib.smoke_simulation_get_view_cuda(self.handle, C.byref(req), C.byref(v)), f"get_view {kind}")
        self.hip.synchronize()
        shape = (v.nz, v.ny, v.nx)
        if v.layout == LAYOUT_F32_3D:
            return self.hip.download(v.data0, shape)
        return tuple(self.hip.download(p, shape) for p in (v.data0, v.data1, v.data2))

    def set_solver_options(self, **kw):
        o = SolverOptions()
        check(self.lib.smoke_simulation_ext_get_solver_options(self.handle, C.byref(o)), "get_solver_options")
        for k, val in kw.items():
            setattr(o, k, val)
        check(self.lib.smoke_simulation_ext_set_solver_options(self.handle, C.byref(o)), "set_solver_options")

    def stats(self) -> SolverStats:
        s = SolverStats()
        check(self.lib.smoke_simulation_ext_get_solver_stats(self.handle, C.byref(s)), "get_solver_stats")
        return s

    def import_state(self, u=None, v=None, w=None, density=None, temperature=None, pressure=None):
        keep = {k: (None if a is None else np.ascontiguousarray(a, dtype=np.float32))
                for k, a in dict(velocity_x=u, velocity_y=v, velocity_z=w, density=density,
                                 temperature=temperature, pressure=pressure).items()}
        hs = HostState()
        for k, a in keep.items():
            setattr(hs, k, None if a is None else a.ctypes.data)
        check(self.lib.smoke_simulation_ext_import_state(self.handle, C.byref(hs)), "import_state")

    def export_state(self) -> dict:
        cfg = self.cfg
        out = {
            "velocity_x": np.empty(face_shape(cfg, 0), np.float32),
            "velocity_y": np.empty(face_shape(cfg, 1), np.float32),
            "velocity_z": np.empty(face_shape(cfg, 2), np.float32),
            "density": np.empty(cell_shape(cfg), np.float32),
            "temperature": np.empty(cell_shape(cfg), np.float32),
            "pressure": np.empty(cell_shape(cfg), np.float32),
            "velocity_star_x": np.empty(face_shape(cfg, 0), np.float32),
            "velocity_star_y": np.empty(face_shape(cfg, 1), np.float32),
            "velocity_star_z": np.empty(face_shape(cfg, 2), np.float32),
        }
        hs = HostState()
        for k, a in out.items():
            setattr(hs, k, a.ctypes.data)
        check(self.lib.smoke_simulation_ext_export_state(self.handle, C.byref(hs)), "export_state")
        return out

    def export_pressure_f64(self) -> np.ndarray:
        out = np.empty(cell_shape(self.cfg), np.float64)
        check(self.lib.smoke_simulation_ext_export_pressure_f64(self.handle, C.c_void_p(out.ctypes.data)), "export_pressure_f64")
        return out

    def plume_sources(self, src: PlumeSource, step: int):
        check(self.lib.smoke_simulation_ext_plume_sources(self.handle, C.byref(src), step), "plume_sources")

    def set_kernel_timing(self, enable: bool):
        check(self.lib.smoke_simulation_ext_set_kernel_timing(self.handle, 1 if enable else 0), "set_kernel_timing")

    def kernel_timing(self) -> KernelTiming:
        t = KernelTiming()
        check(self.lib.smoke_simulation_ext_get_kernel_timing(self.handle, C.byref(t)), "get_kernel_timing")
        return t

    def autotune_launches(self) -> LaunchReport:
        r = LaunchReport()
        check(self.lib.smoke_simulation_ext_autotune_launches(self.handle, C.byref(r)), "autotune_launches")
        return r

    def reset_launches(self):
        check(self.lib.smoke_simulation_ext_reset_launches(self.handle), "reset_launches")

    def write_training_crops(self, path, origins, size, step=0):
        """Appends crops (lower corners `origins`, extent `size` = (sx, sy, sz))
        of the last solve's (b, x, flags) to `path`; see training_data.read."""
        o = np.ascontiguousarray(np.asarray(origins, np.int32).reshape(-1, 3))
        d = TrainingCrops(str(path).encode(), o.ctypes.data, o.shape[0], (C.c_uint32 * 3)(*size), step)
        check(self.lib.smoke_simulation_ext_write_training_crops(self.handle, C.byref(d)), "write_training_crops")

    def copy_training_crops(self, origins, size, d_r: int, d_flags: int):
        """Crops (lower corners `origins`, extent `size` = (sx, sy, sz)) of the
        last solve's unit-RMS right-hand side (FP32) and flags into device
        buffers d_r / d_flags (count * N elements each)."""
        o = np.ascontiguousarray(np.asarray(origins, np.int32).reshape(-1, 3))
        d = DeviceCrops(o.ctypes.data, o.shape[0], (C.c_uint32 * 3)(*size), d_r, d_flags)
        check(self.lib.smoke_simulation_ext_copy_training_crops(self.handle, C.byref(d)), "copy_training_crops")

    def set_learned_preconditioner(self, network):
        """Attaches a learned preconditioner (a LearnedPreconditioner or None)."""
        self._network = network
        check(self.lib.smoke_simulation_ext_set_learned_preconditioner(self.handle, network.handle if network else None),
              "set_learned_preconditioner")

    def launches(self) -> LaunchReport:
        r = LaunchReport()
        check(self.lib.smoke_simulation_ext_get_launches(self.handle, C.byref(r)), "get_launches")
        return r

    def set_launches(self, report: LaunchReport):
        """Applies the launch shapes of `report` (see launches())."""
        check(self.lib.smoke_simulation_ext_set_launches(self.handle, C.byref(report)), "set_launches")

    def set_force_pass(self, fused: bool):
        """Force pass as the fused LDS kernel (default) or as the vorticity and
        cell-force kernels with omega arrays in HBM; the same bits."""
        check(self.lib.smoke_simulation_ext_set_force_pass(self.handle, 1 if fused else 0), "set_force_pass")

    def set_multilevel_options(self, **kw):
        o = MultilevelOptions()
        check(self.lib.smoke_simulation_ext_get_multilevel_options(self.handle, C.byref(o)), "get_multilevel_options")
        for k, val in kw.items():
            setattr(o, k, val)
        check(self.lib.smoke_simulation_ext_set_multilevel_options(self.handle, C.byref(o)), "set_multilevel_options")

    def multilevel_options(self) -> MultilevelOptions:
        o = MultilevelOptions()
        check(self.lib.smoke_simulation_ext_get_multilevel_options(self.handle, C.byref(o)), "get_multilevel_options")
        return o

    def multilevel_info(self) -> tuple[list[tuple[int, int, int]], int]:
        """(dims (nx, ny, nz) of each level, coarsest right-hand sides captured)"""
        i = MultilevelInfo()
        check(self.lib.smoke_simulation_ext_get_multilevel_info(self.handle, C.byref(i)), "get_multilevel_info")
        return [tuple(i.dims[l]) for l in range(i.levels)], int(i.captured)

    def apply_multilevel(self, d_r: int, d_z: int):
        """z = M(r) on device vectors of the pressure grid."""
        check(self.lib.smoke_simulation_ext_apply_multilevel(self.handle, d_r, d_z), "apply_multilevel")

    def level_flags(self, level: int) -> np.ndarray:
        """The flag byte of one level, as a (nz, ny, nx) array."""
        dims, _ = self.multilevel_info()
        nx, ny, nz = dims[level]
        d = self.hip.malloc(nx * ny * nz)
        try:
            check(self.lib.smoke_simulation_ext_export_level_flags(self.handle, level, d), "export_level_flags")
            self.synchronize()
            return self.hip.download(d, (nz, ny, nx), np.uint8)
        finally:
            self.hip.free(d)

    def time_multilevel(self, repetitions: int = 20) -> MultilevelTiming:
        t = MultilevelTiming()
        check(self.lib.smoke_simulation_ext_time_multilevel(self.handle, repetitions, C.byref(t)), "time_multilevel")
        return t

    def set_multilevel_capture(self, d_buffer: int | None, capacity: int):
        check(self.lib.smoke_simulation_ext_set_multilevel_capture(self.handle, d_buffer, capacity), "set_multilevel_capture")


class LearnedPreconditioner:
    """The learned pressure preconditioner (visual-simulation-of-smoke-3d-nn.h):
    create / load weights, train on device volumes, apply."""

    def __init__(self, seed: int = 0, path=None):
        self.lib = L = load_library()
        V = C.c_void_p
        L.smoke_nn_preconditioner_create.argtypes = [C.c_uint64, C.POINTER(V)]
        L.smoke_nn_preconditioner_destroy.argtypes = [V]
        L.smoke_nn_preconditioner_parameter_count.argtypes = [V]
        L.smoke_nn_preconditioner_parameter_count.restype = C.c_int64
        L.smoke_nn_preconditioner_get_parameters.argtypes = [V, V]
        L.smoke_nn_preconditioner_set_parameters.argtypes = [V, V]
        L.smoke_nn_preconditioner_save.argtypes = [V, C.c_char_p]
        L.smoke_nn_preconditioner_load.argtypes = [V, C.c_char_p]
        L.smoke_nn_preconditioner_train_step.argtypes = [V, V, V, C.c_int32, C.POINTER(C.c_int32), C.c_float, C.POINTER(C.c_double), V, V]
        L.smoke_nn_preconditioner_apply.argtypes = [V, V, V, C.POINTER(C.c_int32), V, V]
        L.smoke_nn_preconditioner_apply_static.argtypes = [V, V, V, C.POINTER(C.c_int32), V, V]
        L.smoke_nn_preconditioner_set_loss.argtypes = [V, C.c_uint32]
        self.handle = V()
        check(L.smoke_nn_preconditioner_create(seed, C.byref(self.handle)), "preconditioner create")
        if path is not None:
            self.load(path)

    def close(self):
        if self.handle:
            self.lib.smoke_nn_preconditioner_destroy(self.handle)
            self.handle = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def parameter_count(self) -> int:
        return int(self.lib.smoke_nn_preconditioner_parameter_count(self.handle))

    def parameters(self) -> np.ndarray:
        out = np.empty(self.parameter_count(), np.float32)
        check(self.lib.smoke_nn_preconditioner_get_parameters(self.handle, out.ctypes.data), "get parameters")
        return out

    def save(self, path):
        check(self.lib.smoke_nn_preconditioner_save(self.handle, str(path).encode()), "save weights")

    def load(self, path):
        check(self.lib.smoke_nn_preconditioner_load(self.handle, str(path).encode()), "load weights")

    def set_loss(self, loss: int):
        """NN_LOSS_RESIDUAL or NN_LOSS_ENERGY for later train_step calls."""
        check(self.lib.smoke_nn_preconditioner_set_loss(self.handle, loss), "set loss")

    def apply(self, d_r: int, d_flags: int, dims, d_z: int, stream=None):
        """z = M(r) on one device volume of dims (W, H, D)."""
        check(self.lib.smoke_nn_preconditioner_apply(self.handle, d_r, d_flags, (C.c_int32 * 3)(*dims), d_z, stream), "apply")

    def apply_static(self, d_r: int, d_flags: int, dims, d_z: int, stream=None) -> int:
        """The allocation-free z = M(r) (static device workspace); returns the result code."""
        return self.lib.smoke_nn_preconditioner_apply_static(self.handle, d_r, d_flags, (C.c_int32 * 3)(*dims), d_z, stream)

    def train_step(self, d_r: int, d_flags: int, batch: int, dims, lr: float) -> float:
        """One Adam step on `batch` device volumes (dims = (W, H, D)); returns the loss."""
        loss = C.c_double()
        check(self.lib.smoke_nn_preconditioner_train_step(self.handle, d_r, d_flags, batch, (C.c_int32 * 3)(*dims), lr, C.byref(loss),
                                                          None, None), "train step")
        return loss.value


CONV_KERNELS = {"auto": 0, "gather": 1, "staged": 2}


def set_conv_kernels(choice: str):
    """Process-wide convolution kernel choice (smoke_nn_set_conv_kernels)."""
    L = load_library()
    L.smoke_nn_set_conv_kernels.argtypes = [C.c_int32]
    check(L.smoke_nn_set_conv_kernels(CONV_KERNELS[choice]), "set_conv_kernels")


def inspect_checkpoint(path) -> tuple[int, CheckpointInfo]:
    """Host-only validation of a SMOKECK1 file: (result code, info)."""
    info = CheckpointInfo()
    rc = load_library().smoke_simulation_ext_inspect_checkpoint(str(path).encode(), C.byref(info))
    return rc, info


def measure_copy_bandwidth(nbytes: int, reps: int):
    lib = load_library()
    gbps, ms = C.c_double(), C.c_double()
    check(lib.smoke_simulation_ext_measure_copy_bandwidth(nbytes, reps, C.byref(gbps), C.byref(ms)), "measure_copy_bandwidth")
    return gbps.value, ms.value


# ---------------------------------------------------------------- batched plumes
def batch_lanes(cells, lanes: int) -> list[int]:
    """Lane of each plume under the batch's assignment rule (host only)."""
    lib = load_library()
    n = len(cells)
    arr = (C.c_uint64 * max(n, 1))(*[int(c) for c in cells])
    out = (C.c_int32 * max(n, 1))()
    check(lib.smoke_simulation_ext_batch_lanes(arr, n, lanes, out), "batch_lanes")
    return list(out[:n])


class _BatchPlume(SmokeContext):
    """A plume context owned by a SmokeBatch: every context method applies;
    closing it frees only its uploads (the batch destroys the context)."""

    def __init__(self, batch, handle, cfg):
        self.lib = batch.lib
        self.hip = batch.hip
        self.cfg = cfg
        self.handle = handle
        self._owned = []

    def close(self):
        for p in self._owned:
            self.hip.free(p)
        self._owned = []


class SmokeBatch:
    """Independent plumes of uneven sizes stepped together on `lanes` streams."""

    def __init__(self, cfgs, lanes: int = 0, initial_density=0.0, hip: Hip | None = None):
        self.lib = load_library()
        self.hip = hip or Hip()
        n = len(cfgs)
        descs = (CreateDesc * n)()
        for i, cfg in enumerate(cfgs):
            descs[i].config = cfg
            descs[i].stream = None
            descs[i].initial_density = initial_density
            descs[i].initial_temperature = cfg.ambient_temperature
        d = BatchCreateDesc(C.cast(descs, C.c_void_p), n, lanes)
        h = C.c_void_p()
        check(self.lib.smoke_simulation_ext_create_batch(C.byref(d), C.byref(h)), "create_batch")
        self.handle = h
        self.plumes, self.lanes = [], []
        for i, cfg in enumerate(cfgs):
            ctx, lane = C.c_void_p(), C.c_int32()
            check(self.lib.smoke_simulation_ext_batch_context(h, i, C.byref(ctx), C.byref(lane)), "batch_context")
            self.plumes.append(_BatchPlume(self, ctx, cfg))
            self.lanes.append(lane.value)

    def step(self, sources=None, step: int = 0):
        buf = None
        if sources is not None:
            assert len(sources) == len(self.plumes)
            buf = (PlumeSource * len(sources))(*sources)
        check(self.lib.smoke_simulation_ext_step_batch(self.handle, C.cast(buf, C.c_void_p) if buf is not None else None, step), "step_batch")

    def synchronize(self):
        check(self.lib.smoke_simulation_ext_synchronize_batch(self.handle), "synchronize_batch")

    def set_fusion(self, enable: bool):
        check(self.lib.smoke_simulation_ext_set_batch_fusion(self.handle, 1 if enable else 0), "set_batch_fusion")

    def info(self) -> BatchInfo:
        """How the last step ran (fused path, graph, launches per step and per CG iteration)."""
        out = BatchInfo()
        check(self.lib.smoke_simulation_ext_get_batch_info(self.handle, C.byref(out)), "get_batch_info")
        return out

    def close(self):
        if self.handle:
            self.synchronize()
            for p in self.plumes:
                p.close()
            check(self.lib.smoke_simulation_ext_destroy_batch(self.handle), "destroy_batch")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
