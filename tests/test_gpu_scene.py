"""The reference visualizer's scene as a C-ABI caller (cli/scene_plume.hip, the
pattern of vulkan-app/scene_plume.cpp:223-392): a caller-owned stream, the
emitter uploaded with update_density_source / update_temperature_source, a new
user force every step through update_force_cuda, step_cuda, and views exported
by device-to-device copies on the caller's stream.  The same run in the FP64
oracle, given the emitter weights and per-step forces the harness wrote, must
produce the exported density, temperature and cell-centred velocity.

The scene's solver is the reference's: 96 plain CG iterations from zero per
step (FIXED_ITERATIONS), periodic x and z, no-slip floor, outflow top."""
import struct
import subprocess

import numpy as np
import pytest

import oracle as orc
import smoke_abi as abi
from parity import rel_l2

pytestmark = pytest.mark.gpu

EXE = abi.LIB_PATH.parent / "visual-simulation-of-smoke-scene"


@pytest.fixture(scope="module", autouse=True)
def _require_gpu():
    abi.load_library()
    if abi.Hip().device_count() == 0:
        pytest.skip("no GPU")


def scene_config(nx, ny, nz):
    """scene_plume.ixx:34-78."""
    p = (abi.PERIODIC, 0.0, 0.0, 0.0, 0.0)
    flow = [p, p, (abi.NO_SLIP_WALL, 0.0, 0.0, 0.0, 0.0), (abi.OUTFLOW, 0.0, 0.0, 0.0, 0.0), p, p]
    scal = [(abi.SCALAR_PERIODIC, 0.0)] * 2 + [(abi.FIXED_VALUE, 0.0)] * 2 + [(abi.SCALAR_PERIODIC, 0.0)] * 2
    return abi.make_config(nx, ny, nz, cell_size=0.01, dt=float(np.float32(1.0) / np.float32(90.0)), pressure_iterations=96, ambient=0.0,
                           buoyancy_density=0.8, buoyancy_temperature=0.75, vorticity=0.2, cubic=True, flow=flow, density_bc=scal,
                           temperature_bc=scal)


def emitter(nx, ny, nz):
    """Ellipsoid at (0.5, 0.1, 0.5) of the extent, radii (0.12, 0.08, 0.12),
    weight exp(-2.2 r^2), zero outside."""
    h = np.float32(0.01)
    f = np.float32
    z, y, x = np.meshgrid(np.arange(nz, dtype=f), np.arange(ny, dtype=f), np.arange(nx, dtype=f), indexing="ij")
    dx = ((x + f(0.5)) * h - f(nx) * h * f(0.5)) / (f(nx) * h * f(0.12))
    dy = ((y + f(0.5)) * h - f(ny) * h * f(0.1)) / (f(ny) * h * f(0.08))
    dz = ((z + f(0.5)) * h - f(nz) * h * f(0.5)) / (f(nz) * h * f(0.12))
    r2 = dx * dx + dy * dy + dz * dz
    return np.where(r2 > 1, f(0), np.exp(f(-2.2) * r2)).astype(f)


def read_run(path):
    blob = path.read_bytes()
    assert blob[:8] == b"SMOKESC1"
    nx, ny, nz, steps, seed = struct.unpack_from("<5i", blob, 8)
    cells = nx * ny * nz
    off = 28
    weight = np.frombuffer(blob, np.float32, cells, off).reshape(nz, ny, nx)
    off += 4 * cells
    forces = np.frombuffer(blob, np.float32, 3 * steps, off).reshape(steps, 3)
    off += 12 * steps
    fields = np.frombuffer(blob, np.float32, 5 * cells, off).reshape(5, nz, ny, nx)
    assert off + 20 * cells == len(blob)
    return (nx, ny, nz, steps), weight, forces, fields


@pytest.mark.timeout(300)
def test_scene_caller_matches_oracle(tmp_path):
    dims, steps = (24, 32, 24), 8
    out = tmp_path / "scene.bin"
    print(f"scene harness: {dims} {steps} steps", flush=True)
    r = subprocess.run([str(EXE), "--nx", str(dims[0]), "--ny", str(dims[1]), "--nz", str(dims[2]), "--steps", str(steps), "--seed", "7",
                        "--out", str(out)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    (nx, ny, nz, n), w, forces, fields = read_run(out)
    assert (nx, ny, nz, n) == (*dims, steps)
    # the emitter exp(-2.2 r^2) inside the ellipsoid (scene_plume.cpp:262-284), forces of the scene's speed range
    assert np.allclose(w, emitter(nx, ny, nz), rtol=1e-6, atol=0) and (w > 0).sum() > 50
    speed = np.linalg.norm(forces.astype(np.float64), axis=1)
    assert np.all((speed >= 22 * 0.82 * 0.999) & (speed <= 22 * 1.2 * 1.001)), speed
    assert np.all(forces[:, 1] > 0.9 * speed)

    cfg = scene_config(nx, ny, nz)
    s = orc.OracleState(cfg)
    s["density_source"] = np.float32(18.0) * w
    s["temperature_source"] = np.float32(34.0) * w
    for k in range(steps):
        f = forces[k]
        s["force_user_x"], s["force_user_y"], s["force_user_z"] = w * f[0], w * f[1], w * f[2]
        s.step(orc.options(mode=0))
    s.diagnostics()
    errs = {"density": rel_l2(fields[0], s["density"]), "temperature": rel_l2(fields[1], s["temperature"]),
            "velocity_x": rel_l2(fields[2], s["center_x"]), "velocity_y": rel_l2(fields[3], s["center_y"]),
            "velocity_z": rel_l2(fields[4], s["center_z"])}
    print("PARITY scene " + " ".join(f"{k}={v:.3e}" for k, v in errs.items()), flush=True)
    assert np.abs(fields[0]).max() > 0.1 and np.abs(fields[3]).max() > 0.01
    assert max(errs.values()) < 5e-6, errs      # measured <= 4.4e-7
