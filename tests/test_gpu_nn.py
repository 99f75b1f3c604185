"""The bfloat16 matrix-core convolutions against the float64 restatement in
nn_reference.py: forward (FP32 output, relative L2 <= 1e-5 -- the inputs are
bfloat16 and the products exact, so only the FP32 summation differs), the data
gradient (bfloat16 output, <= 8e-3: one bfloat16 rounding) and the weight and
bias gradients (FP32, <= 1e-5), for every mode, odd and even sizes,
partial 16-voxel tiles, and volumes on both sides of the split-K threshold."""
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
    return L, hip


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
    # 8192+ output tiles: one wave per tile (smaller volumes split K over a block)
    (ref.SAME, (64, 64, 32), 16, 32),
    (ref.UP, (32, 32, 16), 32, 16),
]


@pytest.mark.parametrize("mode,size,cin,cout", CASES)
def test_conv_forward_and_gradients(env, mode, size, cin, cout):
    L, hip = env
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
