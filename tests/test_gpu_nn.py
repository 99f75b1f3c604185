"""The bfloat16 matrix-core convolutions against the float64 restatement in
nn_reference.py: forward (FP32 output, relative L2 <= 1e-5 -- the inputs are
bfloat16 and the products exact, so only the FP32 summation differs), the data
gradient (bfloat16 output, <= 8e-3: one bfloat16 rounding) and the weight and
bias gradients (FP32, <= 1e-5), for every mode, odd and even sizes,
partial 16-voxel tiles, and volumes on both sides of the split-K threshold --
with the automatic kernel choice, and with the LDS-staged kernels forced for
every shape of their table (smoke_nn_set_conv_kernels).  Also: forward and
data gradients of the staged kernels are bitwise those of the gather
kernels, and their weight gradients agree to FP32 summation order."""
import ctypes as C

import numpy as np
import pytest

import nn_reference as ref
import smoke_abi as abi
import training_data as td

pytestmark = pytest.mark.gpu


class Desc(C.Structure):
    _fields_ = [("batch", C.c_int32), ("in_", C.c_int32 * 3), ("out", C.c_int32 * 3), ("cin", C.c_int32), ("cout", C.c_int32),
                ("mode", C.c_int32)]


@pytest.fixture(scope="module")
def env():
    L = abi.load_library()
    hip = abi.Hip()
    if hip.device_count() == 0:
        pytest.skip("no GPU")
    V = C.c_void_p
    L.smoke_nn_packed_weight_bytes.argtypes = [C.POINTER(Desc)]
    L.smoke_nn_packed_weight_bytes.restype = C.c_uint64
    L.smoke_nn_pack_weights.argtypes = [C.POINTER(Desc), V, V, V, V]
    L.smoke_nn_conv_forward.argtypes = [C.POINTER(Desc), V, V, V, V, C.c_int32, V, V, C.c_int32, V]
    L.smoke_nn_conv_backward_data.argtypes = [C.POINTER(Desc), V, V, V, V, V, V]
    L.smoke_nn_conv_backward_weights.argtypes = [C.POINTER(Desc), V, V, V, V, V]
    L.smoke_nn_set_conv_kernels.argtypes = [C.c_int32]
    yield L, hip
    L.smoke_nn_set_conv_kernels(AUTO)


AUTO, GATHER, STAGED = 0, 1, 2


def bf16(a):
    return td.bf16_to_float32(td.float32_to_bf16(np.asarray(a, np.float32))).astype(np.float64)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


CASES = [
    # mode, input (W, H, D), cin, cout
    (ref.SAME, (20, 7, 5), 8, 16),
    (ref.SAME, (16, 6, 4), 16, 32),
    (ref.SAME, (33, 5, 3), 32, 64),
    (ref.SAME, (16, 4, 4), 64, 16),
    (ref.DOWN, (21, 8, 7), 16, 32),
    (ref.DOWN, (16, 6, 6), 32, 64),
    (ref.UP, (10, 4, 3), 64, 32),
    (ref.UP, (9, 5, 4), 32, 16),
    # the remaining shapes of the staged kernels' table
    (ref.SAME, (20, 9, 5), 16, 16),
    (ref.SAME, (18, 7, 4), 32, 32),
    (ref.SAME, (17, 6, 3), 64, 64),
    # 8192+ output tiles: one wave per tile or the staged kernels (smaller
    # volumes split K over a block)
    (ref.SAME, (64, 64, 32), 16, 32),
    (ref.UP, (32, 32, 16), 32, 16),
]


@pytest.mark.parametrize("choice", [AUTO, STAGED])
@pytest.mark.parametrize("mode,size,cin,cout", CASES)
def test_conv_forward_and_gradients(env, mode, size, cin, cout, choice):
    L, hip = env
    assert L.smoke_nn_set_conv_kernels(choice) == 0
    print(f"conv case: mode {mode} size {size} cin {cin} cout {cout} kernels {choice}", flush=True)
    rng = np.random.default_rng(cin * 7 + cout + mode)
    B = 2
    W, H, D = size
    out = tuple(ref.out_size(mode, s) for s in size)
    Wo, Ho, Do = out
    x = bf16(rng.normal(size=(B, D, H, W, cin)))
    w = rng.normal(size=(27, cin, cout)) / np.sqrt(27 * cin)
    wb = bf16(w)
    bias = rng.normal(size=cout)
    skip = bf16(rng.normal(size=(B, Do, Ho, Wo, cout)))
    dy = bf16(rng.normal(size=(B, Do, Ho, Wo, cout)))
    add = bf16(rng.normal(size=(B, D, H, W, cin)))
    mask = bf16(rng.normal(size=(B, D, H, W, cin)))
    desc = Desc(B, (C.c_int32 * 3)(*size), (C.c_int32 * 3)(*out), cin, cout, mode)
    u16 = lambda a: td.float32_to_bf16(np.asarray(a, np.float32))
    ptrs = []

    def up(a):
        p = hip.upload(np.ascontiguousarray(a))
        ptrs.append(p)
        return p

    def alloc(n):
        p = hip.malloc(n)
        hip.lib.hipMemset(C.c_void_p(p), 0, n)
        ptrs.append(p)
        return p

    try:
        pb = int(L.smoke_nn_packed_weight_bytes(C.byref(desc)))
        d_w = up(w.astype(np.float32))
        d_fwd, d_bwd = alloc(pb), alloc(pb)
        assert L.smoke_nn_pack_weights(C.byref(desc), d_w, d_fwd, d_bwd, None) == 0
        d_x, d_bias, d_skip = up(u16(x)), up(bias.astype(np.float32)), up(u16(skip))
        ny = B * Do * Ho * Wo * cout
        d_y, d_yf = alloc(ny * 2), alloc(ny * 4)
        assert L.smoke_nn_conv_forward(C.byref(desc), d_x, d_fwd, d_bias, d_skip, 1, d_y, d_yf, cout, None) == 0
        d_dy, d_add, d_mask = up(u16(dy)), up(u16(add)), up(u16(mask))
        d_dx = alloc(x.size * 2)
        if cin % 16 == 0:
            assert L.smoke_nn_conv_backward_data(C.byref(desc), d_dy, d_bwd, d_add, d_mask, d_dx, None) == 0
        d_dw, d_db = alloc(w.size * 4), alloc(cout * 4)
        assert L.smoke_nn_conv_backward_weights(C.byref(desc), d_x, d_dy, d_dw, d_db, None) == 0
        hip.synchronize()
        yf = hip.download(d_yf, (B, Do, Ho, Wo, cout))
        y16 = td.bf16_to_float32(hip.download(d_y, (B, Do, Ho, Wo, cout), np.uint16))
        dw = hip.download(d_dw, w.shape)
        db = hip.download(d_db, (cout,))
        dx = td.bf16_to_float32(hip.download(d_dx, x.shape, np.uint16)) if cin % 16 == 0 else None
    finally:
        for p in ptrs:
            hip.free(p)

    y_ref = ref.forward(x, wb, bias.astype(np.float32).astype(np.float64), mode, (Do, Ho, Wo), skip=skip, relu=True)
    assert rel(yf, y_ref) <= 1e-5
    np.testing.assert_array_equal(y16, td.bf16_to_float32(td.float32_to_bf16(yf)))
    dw_ref, db_ref = ref.backward_weights(x, dy, mode)
    assert rel(dw, dw_ref) <= 1e-5
    assert rel(db, db_ref) <= 1e-5
    if dx is not None:
        dx_ref = (ref.backward_data(dy, wb, mode, x.shape) + add) * (mask > 0)
        assert rel(dx, dx_ref) <= 8e-3


# Every shape of the staged kernels' table (forward layers and their data
# gradients), on volumes with partial x tiles, partial row groups and a batch.
STAGED_CASES = [
    (ref.SAME, (40, 10, 5), 8, 16),
    (ref.SAME, (72, 10, 5), 16, 16),
    (ref.SAME, (40, 10, 5), 32, 32),
    (ref.SAME, (24, 6, 5), 64, 64),
    (ref.DOWN, (45, 14, 7), 16, 32),
    (ref.DOWN, (40, 10, 6), 32, 64),
    (ref.UP, (20, 5, 4), 64, 32),
    (ref.UP, (36, 7, 5), 32, 16),
]


def _run_conv(L, hip, mode, size, cin, cout, seed, B):
    """Forward (FP32 and bfloat16 outputs), data and weight gradients of one
    convolution with the current kernel choice, as numpy arrays."""
    rng = np.random.default_rng(seed)
    W, H, D = size
    out = tuple(ref.out_size(mode, s) for s in size)
    Wo, Ho, Do = out
    u16 = lambda a: td.float32_to_bf16(np.asarray(a, np.float32))
    x = u16(rng.normal(size=(B, D, H, W, cin)))
    w = (rng.normal(size=(27, cin, cout)) / np.sqrt(27 * cin)).astype(np.float32)
    bias = rng.normal(size=cout).astype(np.float32)
    skip = u16(rng.normal(size=(B, Do, Ho, Wo, cout)))
    dy = u16(rng.normal(size=(B, Do, Ho, Wo, cout)))
    add = u16(rng.normal(size=(B, D, H, W, cin)))
    mask = u16(rng.normal(size=(B, D, H, W, cin)))
    desc = Desc(B, (C.c_int32 * 3)(*size), (C.c_int32 * 3)(*out), cin, cout, mode)
    ptrs = []

    def up(a):
        p = hip.upload(np.ascontiguousarray(a))
        ptrs.append(p)
        return p

    def alloc(n):
        p = hip.malloc(n)
        hip.lib.hipMemset(C.c_void_p(p), 0, n)
        ptrs.append(p)
        return p

    try:
        pb = int(L.smoke_nn_packed_weight_bytes(C.byref(desc)))
        d_fwd, d_bwd = alloc(pb), alloc(pb)
        assert L.smoke_nn_pack_weights(C.byref(desc), up(w), d_fwd, d_bwd, None) == 0
        d_x = up(x)
        ny = B * Do * Ho * Wo * cout
        d_y, d_yf = alloc(ny * 2), alloc(ny * 4)
        assert L.smoke_nn_conv_forward(C.byref(desc), d_x, d_fwd, up(bias), up(skip), 1, d_y, d_yf, cout, None) == 0
        d_dy = up(dy)
        d_dx = alloc(x.size * 2)
        if cin % 16 == 0:
            assert L.smoke_nn_conv_backward_data(C.byref(desc), d_dy, d_bwd, up(add), up(mask), d_dx, None) == 0
        d_dw, d_db = alloc(w.size * 4), alloc(cout * 4)
        assert L.smoke_nn_conv_backward_weights(C.byref(desc), d_x, d_dy, d_dw, d_db, None) == 0
        hip.synchronize()
        return {"yf": hip.download(d_yf, (B, Do, Ho, Wo, cout)), "y": hip.download(d_y, (B, Do, Ho, Wo, cout), np.uint16),
                "dx": hip.download(d_dx, x.shape, np.uint16), "dw": hip.download(d_dw, w.shape), "db": hip.download(d_db, (cout,))}
    finally:
        for p in ptrs:
            hip.free(p)


@pytest.mark.parametrize("mode,size,cin,cout", STAGED_CASES)
def test_staged_kernels_equal_gather_kernels(env, mode, size, cin, cout):
    """The gather kernel runs one wave per 16-voxel tile from 8192 tiles on
    (smaller volumes split K over a block, which sums in another order): the
    batch is chosen so that the forward output and the data gradient's output
    (the input volume) both reach that."""
    L, hip = env
    out = [ref.out_size(mode, s) for s in size]
    per_volume = min(out[2] * out[1] * -(-out[0] // 16), size[2] * size[1] * -(-size[0] // 16))
    B = -(-8192 // per_volume)
    print(f"staged vs gather: mode {mode} size {size} cin {cin} cout {cout} batch {B}", flush=True)
    try:
        assert L.smoke_nn_set_conv_kernels(GATHER) == 0
        g = _run_conv(L, hip, mode, size, cin, cout, seed=cin + 3 * cout + mode, B=B)
        assert L.smoke_nn_set_conv_kernels(STAGED) == 0
        s = _run_conv(L, hip, mode, size, cin, cout, seed=cin + 3 * cout + mode, B=B)
    finally:
        L.smoke_nn_set_conv_kernels(AUTO)
    np.testing.assert_array_equal(s["yf"], g["yf"])
    np.testing.assert_array_equal(s["y"], g["y"])
    np.testing.assert_array_equal(s["dx"], g["dx"])
    e = rel(s["dw"].astype(np.float64), g["dw"].astype(np.float64))
    print(f"PARITY staged_wgrad mode={mode} cin={cin} cout={cout} rel={e:.3e}", flush=True)
    assert e <= 2e-6
    np.testing.assert_allclose(s["db"], g["db"], rtol=1e-5, atol=1e-5 * np.abs(g["db"]).max())


def test_conv_kernel_choice_rejects_unknown_values(env):
    L, _ = env
    assert L.smoke_nn_set_conv_kernels(3) != 0
    assert L.smoke_nn_set_conv_kernels(-1) != 0
    assert L.smoke_nn_set_conv_kernels(AUTO) == 0
