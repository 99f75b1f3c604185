"""ctypes binding of the CPU oracle (oracle/smoke_oracle.cpp) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product library never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
REPO = ORACLE_DIR.parent
LIB = ORACLE_DIR / "liboracle.so"
sys.path.insert(0, str(REPO / "visual-simulation-of-smoke_amd"))
import smoke_abi as abi  # noqa: E402  (struct layouts of the C ABI)


def build(verbose: bool = False) -> Path:
    """Compiles the oracle with g++ (OpenMP, -O2, no fast-math so FP64 stays IEEE)."""
    cmd = ["g++", "-O2", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-ffp-contract=off",
           "-Wall", "-Wextra", "-o", str(LIB) + ".tmp", str(ORACLE_DIR / "smoke_oracle.cpp")]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(str(LIB) + ".tmp", LIB)
    return LIB


class Options(C.Structure):
    _fields_ = [("pressure_mode", C.c_int32), ("preconditioner", C.c_int32), ("relative_tolerance", C.c_double),
                ("max_iterations", C.c_int32), ("max_refinements", C.c_int32), ("warm_start", C.c_int32),
                ("adaptive_dt", C.c_int32), ("cfl", C.c_double), ("dt_max", C.c_double),
                ("inner_reduction", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("refinements", C.c_int32), ("converged", C.c_int32),
                ("anchor", C.c_int32), ("relative_residual", C.c_double), ("rhs_norm", C.c_double),
                ("dt", C.c_double), ("max_face_speed", C.c_double)]


STATE_FIELDS = [
    "u", "v", "w", "us", "vs", "ws", "density", "temperature", "density_source", "temperature_source",
    "force_user_x", "force_user_y", "force_user_z", "solid_velocity_x", "solid_velocity_y", "solid_velocity_z",
    "occupancy", "solid_temperature", "pressure", "pressure_rhs", "force_x", "force_y", "force_z",
    "center_x", "center_y", "center_z", "vorticity_x", "vorticity_y", "vorticity_z", "vorticity_magnitude",
    "divergence", "velocity_magnitude",
]


class _State(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in STATE_FIELDS]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        P = C.c_void_p
        D = C.c_double
        cfgp = C.POINTER(abi.Config)
        L.oracle_run_stages.argtypes = [cfgp, C.POINTER(Options), C.POINTER(_State), C.c_int, C.c_int, C.POINTER(Stats)]
        L.oracle_diagnostics.argtypes = [cfgp, C.POINTER(_State)]
        L.oracle_monotonic_cubic.argtypes = [D, D, D, D, D]
        L.oracle_monotonic_cubic.restype = D
        L.oracle_sample.argtypes = [cfgp, C.c_int, C.c_int, P, D, D, D]
        L.oracle_sample.restype = D
        L.oracle_load.argtypes = [cfgp, C.c_int, P, C.c_int, C.c_int, C.c_int]
        L.oracle_load.restype = D
        L.oracle_load_center_component.argtypes = [cfgp, C.c_int, P, C.c_int, C.c_int, C.c_int]
        L.oracle_load_center_component.restype = D
        L.oracle_trace.argtypes = [cfgp, P, P, P, P, D, P, P]
        L.oracle_pressure_anchor.argtypes = [cfgp, P]
        L.oracle_pressure_apply.argtypes = [cfgp, P, P, P]
        L.oracle_pressure_dense.argtypes = [cfgp, P, P]
        L.oracle_pressure_rhs.argtypes = [cfgp, P, P, P, P, D, P]
        L.oracle_nearest_fluid.argtypes = [cfgp, P, P]
        L.oracle_plume_sources.argtypes = [cfgp, C.POINTER(abi.PlumeSource), C.c_uint32, P, P]
        L.oracle_noise_bits.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.oracle_noise_bits.restype = C.c_uint32
        L.oracle_f32_benchmark.argtypes = [cfgp, C.POINTER(Options), C.POINTER(abi.PlumeSource), C.c_int, C.c_int,
                                           C.c_int, C.POINTER(C.c_double), C.POINTER(Stats)]
        L.oracle_f32_steps.argtypes = [cfgp, C.POINTER(Options), C.POINTER(abi.PlumeSource), P, C.c_int, C.c_int, C.c_int,
                                       P, P, C.POINTER(Stats)]
        L.oracle_f32_step_sample.argtypes = [cfgp, C.POINTER(Options), C.POINTER(abi.PlumeSource), C.c_int, C.c_int,
                                             C.POINTER(C.c_double), C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _p(a):
    return C.c_void_p(a.ctypes.data)


def options(mode=1, preconditioner=1, tol=1e-6, max_iterations=20000, max_refinements=8, warm_start=0,
            adaptive_dt=0, cfl=0.5, dt_max=0.01, inner_reduction=1e-12) -> Options:
    return Options(mode, preconditioner, tol, max_iterations, max_refinements, warm_start, adaptive_dt, cfl, dt_max,
                   inner_reduction)


class OracleState:
    """FP64 state of the oracle, laid out like the GPU state (numpy z,y,x)."""

    def __init__(self, cfg: abi.Config, initial_density=0.0, initial_temperature=None):
        self.cfg = cfg
        cs = abi.cell_shape(cfg)
        t0 = cfg.ambient_temperature if initial_temperature is None else initial_temperature
        self.a = {}
        for n in STATE_FIELDS:
            if n in ("u", "us"):
                shape = abi.face_shape(cfg, 0)
            elif n in ("v", "vs"):
                shape = abi.face_shape(cfg, 1)
            elif n in ("w", "ws"):
                shape = abi.face_shape(cfg, 2)
            else:
                shape = cs
            self.a[n] = np.zeros(shape, np.uint8 if n == "occupancy" else np.float64)
        self.a["density"][:] = initial_density
        self.a["temperature"][:] = t0
        self.a["solid_temperature"][:] = cfg.ambient_temperature
        self.stats = Stats()
        self.stats.dt = cfg.dt

    def __getitem__(self, k):
        return self.a[k]

    def __setitem__(self, k, v):
        self.a[k][...] = v

    def _struct(self):
        s = _State()
        for n in STATE_FIELDS:
            setattr(s, n, self.a[n].ctypes.data)
        return s

    def run_stages(self, first: int, last: int, opts: Options | None = None) -> Stats:
        opts = opts or options()
        s = self._struct()
        lib().oracle_run_stages(C.byref(self.cfg), C.byref(opts), C.byref(s), first, last, C.byref(self.stats))
        return self.stats

    def step(self, opts: Options | None = None) -> Stats:
        return self.run_stages(0, 4, opts)

    def diagnostics(self):
        s = self._struct()
        lib().oracle_diagnostics(C.byref(self.cfg), C.byref(s))


def monotonic_cubic(p0, p1, p2, p3, t):
    return lib().oracle_monotonic_cubic(p0, p1, p2, p3, t)


FIELD_DENSITY, FIELD_TEMPERATURE, FIELD_U, FIELD_V, FIELD_W, FIELD_FLOW_CELL = 0, 1, 2, 3, 4, 5


def sample(cfg, field_kind, cubic, field, x, y, z):
    f = np.ascontiguousarray(field, np.float64)
    return lib().oracle_sample(C.byref(cfg), field_kind, int(cubic), _p(f), x, y, z)


def load(cfg, field_kind, field, i, j, k):
    f = np.ascontiguousarray(field, np.float64)
    return lib().oracle_load(C.byref(cfg), field_kind, _p(f), i, j, k)


def load_center_component(cfg, comp, field, x, y, z):
    f = np.ascontiguousarray(field, np.float64)
    return lib().oracle_load_center_component(C.byref(cfg), comp, _p(f), x, y, z)


def trace(cfg, u, v, w, occ, dt, start):
    u, v, w = (np.ascontiguousarray(a, np.float64) for a in (u, v, w))
    occ = np.ascontiguousarray(occ, np.uint8)
    s = np.asarray(start, np.float64)
    out = np.zeros(3)
    lib().oracle_trace(C.byref(cfg), _p(u), _p(v), _p(w), _p(occ), dt, _p(s), _p(out))
    return out


def pressure_anchor(cfg, occ):
    occ = np.ascontiguousarray(occ, np.uint8)
    return lib().oracle_pressure_anchor(C.byref(cfg), _p(occ))


def pressure_apply(cfg, occ, x):
    occ = np.ascontiguousarray(occ, np.uint8)
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros_like(x)
    lib().oracle_pressure_apply(C.byref(cfg), _p(occ), _p(x), _p(y))
    return y


def pressure_dense(cfg, occ):
    occ = np.ascontiguousarray(occ, np.uint8)
    n = occ.size
    a = np.zeros((n, n))
    lib().oracle_pressure_dense(C.byref(cfg), _p(occ), _p(a))
    return a


def pressure_rhs(cfg, occ, us, vs, ws, dt):
    occ = np.ascontiguousarray(occ, np.uint8)
    us, vs, ws = (np.ascontiguousarray(a, np.float64) for a in (us, vs, ws))
    rhs = np.zeros(occ.shape)
    lib().oracle_pressure_rhs(C.byref(cfg), _p(occ), _p(us), _p(vs), _p(ws), dt, _p(rhs))
    return rhs


def nearest_fluid(cfg, occ):
    occ = np.ascontiguousarray(occ, np.uint8)
    idx = np.zeros(occ.shape, np.int64)
    lib().oracle_nearest_fluid(C.byref(cfg), _p(occ), _p(idx))
    return idx


def plume_sources(cfg, src: abi.PlumeSource, step: int):
    rs = np.zeros(abi.cell_shape(cfg))
    ts = np.zeros(abi.cell_shape(cfg))
    lib().oracle_plume_sources(C.byref(cfg), C.byref(src), step, _p(rs), _p(ts))
    return rs, ts


def noise_bits(seed, step, cell):
    return lib().oracle_noise_bits(seed, step, cell)


def f32_benchmark(cfg, opts: Options, src: abi.PlumeSource | None, warmup: int, steps: int, threads: int):
    secs = C.c_double()
    st = Stats()
    lib().oracle_f32_benchmark(C.byref(cfg), C.byref(opts), C.byref(src) if src is not None else None, warmup, steps,
                               threads, C.byref(secs), C.byref(st))
    return secs.value, st


def f32_steps(cfg, opts: Options, src: abi.PlumeSource | None, occupancy, warmup: int, steps: int, threads: int):
    """Whole FP32 host-port steps: (seconds per timed step, pressure iterations per timed step)."""
    secs = np.zeros(max(steps, 1))
    its = np.zeros(max(steps, 1), np.int32)
    occ = None if occupancy is None else np.ascontiguousarray(occupancy, np.uint8)
    lib().oracle_f32_steps(C.byref(cfg), C.byref(opts), C.byref(src) if src is not None else None,
                           None if occ is None else _p(occ), warmup, steps, threads, _p(secs), _p(its), C.byref(Stats()))
    return secs[:steps].tolist(), its[:steps].tolist()


def f32_step_sample(cfg, opts: Options, src: abi.PlumeSource | None, cg_iterations: int, threads: int):
    """(seconds of one step without its pressure iterations, seconds per CG iteration)."""
    other, it = C.c_double(), C.c_double()
    lib().oracle_f32_step_sample(C.byref(cfg), C.byref(opts), C.byref(src) if src is not None else None, cg_iterations, threads,
                                 C.byref(other), C.byref(it))
    return other.value, it.value
