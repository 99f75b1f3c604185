"""Training-pair crops written by the HIP library, read back with
training_data.read and checked bitwise against the solve they came from:
scale = RMS of b over the crop, r = bf16(b / scale), z = bf16(x / scale)
with x the FP64 solution, and the operator flags (identity rows for solids
and the anchor, fluid-neighbour bits otherwise).  The device-resident copy
of the same crops (smoke_simulation_ext_copy_training_crops) holds FP32
b / scale and the same flags."""
import numpy as np
import pytest

import oracle as orc
import smoke_abi as abi
import training_data as td
from parity import load_gpu, make_state, tolerance_options

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _require_gpu():
    abi.load_library()
    if abi.Hip().device_count() == 0:
        pytest.skip("no GPU")


def expected_flags(occ, anchor):
    nz, ny, nx = occ.shape
    fluid = occ == 0
    f = np.zeros(occ.shape, np.uint8)
    pad = np.zeros((nz + 2, ny + 2, nx + 2), bool)
    pad[1:-1, 1:-1, 1:-1] = fluid
    shifts = [(0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)]   # x-, x+, y-, y+, z-, z+
    for bit, (dz, dy, dx) in enumerate(shifts):
        nb = pad[1 + dz:nz + 1 + dz, 1 + dy:ny + 1 + dy, 1 + dx:nx + 1 + dx]
        f |= (nb.astype(np.uint8) << bit)
    ident = ~fluid
    ident.ravel()[anchor] = True
    f[ident] = 0x40
    return f


def test_training_crops_roundtrip(tmp_path):
    cfg = abi.make_config(40, 48, 36, cell_size=1.0 / 48, dt=0.01, vorticity=2.0, cubic=True)
    st = make_state(cfg, seed=3, solids=2)
    origins = [(0, 0, 0), (24, 28, 24), (10, 5, 7)]
    size = (16, 20, 12)
    path = tmp_path / "pairs.bin"
    hip = abi.Hip()
    cells = size[0] * size[1] * size[2]
    d_r, d_f = hip.malloc(4 * cells * 3), hip.malloc(cells * 3)
    try:
        with abi.SmokeContext(cfg, hip=hip) as ctx:
            load_gpu(ctx, st)
            tolerance_options(ctx)
            ctx.step(2)
            ctx.write_training_crops(path, origins[:2], size, step=2)
            ctx.write_training_crops(path, origins[2:], size, step=2)   # appends
            # the same crops kept on the device (FP32 r)
            ctx.copy_training_crops(origins, size, d_r, d_f)
            ctx.synchronize()
            b = ctx.view("FLOW_PRESSURE_RHS")
            x = ctx.export_pressure_f64()
        dev_r = hip.download(d_r, (3, size[2], size[1], size[0]))
        dev_f = hip.download(d_f, (3, size[2], size[1], size[0]), np.uint8)
    finally:
        hip.free(d_r)
        hip.free(d_f)
    d = td.read(path)
    man = td.read_manifest(path)
    assert d["crop"] == size and d["grid"] == (40, 48, 36) and len(d["scale"]) == 3
    assert [m["offset"] for m in man] == [64 + k * (32 + 5 * 16 * 20 * 12) for k in range(3)]
    occ = st["occupancy"]
    flags = expected_flags(occ, orc.pressure_anchor(cfg, occ))
    for k, (ox, oy, oz) in enumerate(origins):
        sl = np.s_[oz:oz + size[2], oy:oy + size[1], ox:ox + size[0]]
        bc = b[sl].astype(np.float64)
        assert tuple(d["origin"][k]) == (ox, oy, oz) and d["step"][k] == 2
        scale = np.sqrt(np.mean(bc * bc))
        assert abs(d["scale"][k] - scale) <= 1e-12 * scale
        s = d["scale"][k]
        np.testing.assert_array_equal(d["r_bits"][k], td.float32_to_bf16((bc / s).astype(np.float32)))
        np.testing.assert_array_equal(d["z_bits"][k], td.float32_to_bf16((x[sl] / s).astype(np.float32)))
        np.testing.assert_array_equal(d["flags"][k], flags[sl])
        # device crops: FP32 b / scale (the file's r is its bfloat16) and the same flags
        np.testing.assert_array_equal(dev_r[k], (bc / s).astype(np.float32))
        np.testing.assert_array_equal(dev_f[k], flags[sl])
