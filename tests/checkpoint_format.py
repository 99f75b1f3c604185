"""An independent writer/reader of the SMOKECK1 checkpoint format, written
from the layout documented with smoke_simulation_ext_save_checkpoint
(include/visual-simulation-of-smoke-3d-ext.h), with struct.pack and numpy only.

    char[8] "SMOKECK1"; u32 version = 2; i32 nx, ny, nz; u32 count = 18;
    18 x (char[24] name, u64 bytes, data); i32 warm-start count; EOF.
"""
from __future__ import annotations

import struct

import numpy as np

MAGIC = b"SMOKECK1"
VERSION = 2


def fields(nx: int, ny: int, nz: int):
    """(name, numpy dtype, shape z-y-x) of every field, in file order."""
    c = (nz, ny, nx)
    f32, f64 = np.float32, np.float64
    out = [("velocity_x", f32, (nz, ny, nx + 1)), ("velocity_y", f32, (nz, ny + 1, nx)), ("velocity_z", f32, (nz + 1, ny, nx))]
    out += [(n, f32, c) for n in ("density", "temperature", "density_source", "temperature_source", "force_x", "force_y", "force_z",
                                  "solid_velocity_x", "solid_velocity_y", "solid_velocity_z", "solid_temperature")]
    out += [("occupancy", np.uint8, c), ("pressure", f32, c), ("pressure_f64", f64, c), ("pressure_f64_history", f64, c)]
    return out


def random_state(nx: int, ny: int, nz: int, seed: int = 0) -> dict:
    rng = np.random.default_rng(seed)
    st = {}
    for name, dt, shape in fields(nx, ny, nz):
        if dt == np.uint8:
            st[name] = (rng.uniform(size=shape) < 0.2).astype(np.uint8)
        else:
            st[name] = rng.normal(size=shape).astype(dt)
    return st


def assemble(nx: int, ny: int, nz: int, state: dict, solves_done: int = 0) -> bytes:
    parts = [MAGIC, struct.pack("<I3iI", VERSION, nx, ny, nz, len(fields(nx, ny, nz)))]
    for name, dt, shape in fields(nx, ny, nz):
        a = np.ascontiguousarray(state[name], dt)
        assert a.shape == shape, (name, a.shape, shape)
        parts.append(name.encode().ljust(24, b"\0"))
        parts.append(struct.pack("<Q", a.nbytes))
        parts.append(a.astype(np.dtype(dt).newbyteorder("<"), copy=False).tobytes())
    parts.append(struct.pack("<i", solves_done))
    return b"".join(parts)


def parse(blob: bytes):
    """(dims, state, solves_done) of a file; raises ValueError on any mismatch."""
    try:
        return _parse(blob)
    except struct.error as e:
        raise ValueError(f"truncated: {e}") from None


def _parse(blob: bytes):
    if blob[:8] != MAGIC:
        raise ValueError("magic")
    version, nx, ny, nz, count = struct.unpack_from("<I3iI", blob, 8)
    if version != VERSION:
        raise ValueError("version")
    if min(nx, ny, nz) <= 0:
        raise ValueError("dimensions")
    layout = fields(nx, ny, nz)
    if count != len(layout):
        raise ValueError("count")
    off, st = 28, {}
    for name, dt, shape in layout:
        got = blob[off:off + 24].rstrip(b"\0").decode()
        (nbytes,) = struct.unpack_from("<Q", blob, off + 24)
        off += 32
        if got != name or nbytes != int(np.prod(shape)) * np.dtype(dt).itemsize:
            raise ValueError(name)
        if off + nbytes > len(blob):
            raise ValueError("truncated")
        st[name] = np.frombuffer(blob, np.dtype(dt).newbyteorder("<"), int(np.prod(shape)), off).reshape(shape)
        off += nbytes
    (solves_done,) = struct.unpack_from("<i", blob, off)
    if off + 4 != len(blob):
        raise ValueError("trailing")
    if not 0 <= solves_done <= 2:
        raise ValueError("warm-start count")
    return (nx, ny, nz), st, solves_done
