"""The learned preconditioner's training step against the float64
restatement of the whole network (nn_reference.unet_loss_and_gradients,
which rounds every stored tensor to bfloat16 where the GPU does): loss to
1e-3 relative, each layer's weight and bias gradients to 3e-2 relative L2
(FP32 against float64 summation, and the rare bfloat16 rounding that flips
between them).  Also: Adam steps on one batch lower the loss, inference
zeroes identity rows and scales with its input, and weights round-trip
through a file."""
import ctypes as C

import numpy as np
import pytest

import nn_reference as ref
import smoke_abi as abi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    L = abi.load_library()
    hip = abi.Hip()
    if hip.device_count() == 0:
        pytest.skip("no GPU")
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
    return L, hip


def volumes(batch, dims, seed):
    """r and operator flags of crops cut from a larger occupancy with solids
    (so that boundary bits can point outside the crop) and an anchor."""
    W, H, D = dims
    rng = np.random.default_rng(seed)
    occ = np.zeros((batch, D + 2, H + 2, W + 2), bool)
    for b in range(batch):
        z0, y0, x0 = rng.integers(0, [D, H, W])
        occ[b, z0:z0 + 3, y0:y0 + 3, x0:x0 + 4] = True
    fluid = ~occ
    pad = np.pad(fluid, ((0, 0), (1, 1), (1, 1), (1, 1)))
    flags = np.zeros(occ.shape, np.uint8)
    shifts = [(0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)]
    Dp, Hp, Wp = D + 2, H + 2, W + 2
    for bit, (dz, dy, dx) in enumerate(shifts):
        nb = pad[:, 1 + dz:Dp + 1 + dz, 1 + dy:Hp + 1 + dy, 1 + dx:Wp + 1 + dx]
        flags |= (nb.astype(np.uint8) << bit)
    flags[occ] = 0x40
    flags = np.ascontiguousarray(flags[:, 1:-1, 1:-1, 1:-1])
    flags[:, 0, 0, 0] = 0x40   # an anchor-like identity row
    r = rng.normal(size=(batch, D, H, W)) * (1.0 + 0.5 * np.sin(np.arange(W) / 3.0))
    r = np.where(flags & 0x40, 0.0, r).astype(np.float32)
    return r, flags


def test_training_step_matches_float64_network(lib, tmp_path):
    L, hip = lib
    net = C.c_void_p()
    assert L.smoke_nn_preconditioner_create(7, C.byref(net)) == 0
    ptrs = []
    try:
        n = L.smoke_nn_preconditioner_parameter_count(net)
        params = np.empty(n, np.float32)
        assert L.smoke_nn_preconditioner_get_parameters(net, params.ctypes.data) == 0
        batch, dims = 2, (16, 12, 8)
        r, flags = volumes(batch, dims, 3)
        d_r, d_f = hip.upload(r), hip.upload(flags)
        ptrs += [d_r, d_f]
        cd = (C.c_int32 * 3)(*dims)
        loss = C.c_double()
        grads = np.empty(n, np.float32)
        assert L.smoke_nn_preconditioner_train_step(net, d_r, d_f, batch, cd, 0.0, C.byref(loss), grads.ctypes.data, None) == 0
        ref_loss, ref_grads = ref.unet_loss_and_gradients(params, r, flags)
        assert abs(loss.value - ref_loss) <= 1e-3 * ref_loss
        off = 0
        for li, (_, cin, cout, *_) in enumerate(ref.LAYERS):
            for size in (27 * cin * cout, cout):
                g, rg = grads[off:off + size], ref_grads[off:off + size]
                off += size
                norm = np.linalg.norm(rg)
                if norm == 0:
                    assert np.all(g == 0)
                else:
                    assert np.linalg.norm(g - rg) <= 3e-2 * norm, (li, size, np.linalg.norm(g - rg) / norm)
        # Adam on the same batch lowers the loss
        first = loss.value
        for _ in range(30):
            assert L.smoke_nn_preconditioner_train_step(net, d_r, d_f, batch, cd, 1e-3, C.byref(loss), None, None) == 0
        assert loss.value < 0.8 * first, (first, loss.value)
        # save / load round trip
        path = str(tmp_path / "w.bin").encode()
        assert L.smoke_nn_preconditioner_save(net, path) == 0
        trained = np.empty(n, np.float32)
        assert L.smoke_nn_preconditioner_get_parameters(net, trained.ctypes.data) == 0
        net2 = C.c_void_p()
        assert L.smoke_nn_preconditioner_create(99, C.byref(net2)) == 0
        assert L.smoke_nn_preconditioner_load(net2, path) == 0
        back = np.empty(n, np.float32)
        assert L.smoke_nn_preconditioner_get_parameters(net2, back.ctypes.data) == 0
        np.testing.assert_array_equal(trained, back)
        # inference: identity rows give 0, M(a r) = a M(r) up to bfloat16 rounding
        d_z = hip.malloc(r[0].nbytes)
        d_r2 = hip.upload(4.0 * r[0])
        d_z2 = hip.malloc(r[0].nbytes)
        ptrs += [d_z, d_r2, d_z2]
        d_r0 = hip.upload(r[0])
        ptrs.append(d_r0)
        d_f0 = hip.upload(flags[0])
        ptrs.append(d_f0)
        assert L.smoke_nn_preconditioner_apply(net2, d_r0, d_f0, cd, d_z, None) == 0
        assert L.smoke_nn_preconditioner_apply(net2, d_r2, d_f0, cd, d_z2, None) == 0
        hip.synchronize()
        z = hip.download(d_z, r[0].shape)
        z2 = hip.download(d_z2, r[0].shape)
        assert np.all(z[flags[0] & 0x40 != 0] == 0) and np.isfinite(z).all()
        np.testing.assert_allclose(z2, 4.0 * z, rtol=1e-5, atol=1e-6 * np.abs(z).max())
        L.smoke_nn_preconditioner_destroy(net2)
    finally:
        for p in ptrs:
            hip.free(p)
        L.smoke_nn_preconditioner_destroy(net)


def test_energy_loss_training_step_matches_float64_network(lib):
    """The energy loss (smoke_nn_preconditioner_set_loss): loss value and
    every layer's gradients against the float64 restatement."""
    L, hip = lib
    L.smoke_nn_preconditioner_set_loss.argtypes = [C.c_void_p, C.c_uint32]
    L.smoke_nn_preconditioner_set_loss.restype = C.c_int
    net = C.c_void_p()
    assert L.smoke_nn_preconditioner_create(11, C.byref(net)) == 0
    ptrs = []
    try:
        assert L.smoke_nn_preconditioner_set_loss(net, 2) != 0
        assert L.smoke_nn_preconditioner_set_loss(net, 1) == 0
        n = L.smoke_nn_preconditioner_parameter_count(net)
        params = np.empty(n, np.float32)
        assert L.smoke_nn_preconditioner_get_parameters(net, params.ctypes.data) == 0
        batch, dims = 2, (16, 12, 8)
        r, flags = volumes(batch, dims, 5)
        d_r, d_f = hip.upload(r), hip.upload(flags)
        ptrs += [d_r, d_f]
        cd = (C.c_int32 * 3)(*dims)
        loss = C.c_double()
        grads = np.empty(n, np.float32)
        assert L.smoke_nn_preconditioner_train_step(net, d_r, d_f, batch, cd, 0.0, C.byref(loss), grads.ctypes.data, None) == 0
        ref_loss, ref_grads = ref.unet_loss_and_gradients(params, r, flags, loss="energy")
        assert abs(loss.value - ref_loss) <= 1e-3 * abs(ref_loss) + 1e-6, (loss.value, ref_loss)
        off = 0
        for li, (_, cin, cout, *_) in enumerate(ref.LAYERS):
            for size in (27 * cin * cout, cout):
                g, rg = grads[off:off + size], ref_grads[off:off + size]
                off += size
                norm = np.linalg.norm(rg)
                if norm == 0:
                    assert np.all(g == 0)
                else:
                    # 8 %: bfloat16 tie flips between the two summation
                    # orders.  The worst layer is the 4^3 one (64 voxels per
                    # volume): 3.1 % with the K steps of each output tile
                    # split over 4 waves, 5.0 % with them split over 16
                    # (another FP32 order of the same products)
                    assert np.linalg.norm(g - rg) <= 8e-2 * norm, (li, size, np.linalg.norm(g - rg) / norm)
        first = loss.value
        for _ in range(30):
            assert L.smoke_nn_preconditioner_train_step(net, d_r, d_f, batch, cd, 1e-3, C.byref(loss), None, None) == 0
        assert loss.value < first - 0.1 * abs(first), (first, loss.value)
    finally:
        for p in ptrs:
            hip.free(p)
        L.smoke_nn_preconditioner_destroy(net)


@pytest.mark.timeout(300)
def test_training_on_64_cubed_crops_of_real_solves(lib):
    """64^3 crops of the 128^3 plume's Jacobi PCG solves (the training data
    of the 64^3-crop network, kept on the device): one crop's loss and every
    layer's gradients against the float64 network at the shape the network
    trains on, then Adam on a batch of four crops lowers the loss."""
    import train_preconditioner as tp
    L, hip = lib
    n, crop, count = 128, 64, 4
    cells = crop ** 3
    d_r, d_f = hip.malloc(4 * cells * count), hip.malloc(cells * count)
    net = C.c_void_p()
    assert L.smoke_nn_preconditioner_create(11, C.byref(net)) == 0
    try:
        cfg, src = tp.plume_config(n)
        with abi.SmokeContext(cfg, hip=hip) as ctx:
            ctx.set_solver_options(pressure_mode=abi.PRESSURE_TOLERANCE, preconditioner=abi.PRECONDITIONER_JACOBI, relative_tolerance=1e-6,
                                   max_iterations=50000, max_refinements=4, warm_start=2, adaptive_dt=1, cfl=0.5, dt_max=0.01,
                                   check_interval=32)
            for k in range(6):
                ctx.plume_sources(src, k)
                ctx.step()
            # crops around the plume (the source sits at x = z = 64, y = 16)
            origins = [(32, 0, 32), (0, 0, 0), (64, 10, 32), (16, 40, 48)]
            ctx.copy_training_crops(origins, (crop, crop, crop), d_r, d_f)
            ctx.synchronize()
        r = hip.download(d_r, (count, crop, crop, crop))
        flags = hip.download(d_f, (count, crop, crop, crop), np.uint8)
        assert np.isfinite(r).all() and np.abs(r[0]).max() > 0

        npar = L.smoke_nn_preconditioner_parameter_count(net)
        params = np.empty(npar, np.float32)
        assert L.smoke_nn_preconditioner_get_parameters(net, params.ctypes.data) == 0
        cd = (C.c_int32 * 3)(crop, crop, crop)
        loss = C.c_double()
        grads = np.empty(npar, np.float32)
        print("64^3 crop: GPU training step", flush=True)
        assert L.smoke_nn_preconditioner_train_step(net, d_r, d_f, 1, cd, 0.0, C.byref(loss), grads.ctypes.data, None) == 0
        print("64^3 crop: float64 network", flush=True)
        abs_sums = []
        ref_loss, ref_grads = ref.unet_loss_and_gradients(params, r[:1], flags[:1], abs_sums=abs_sums)
        rel_loss = abs(loss.value - ref_loss) / ref_loss
        worst, worst_bias, off = 0.0, 0.0, 0
        for li, (_, cin, cout, *_) in enumerate(ref.LAYERS):
            # weights: relative L2 (FP32 against float64 sums, rare bfloat16 flips)
            size = 27 * cin * cout
            g, rg = grads[off:off + size], ref_grads[off:off + size]
            off += size
            e = np.linalg.norm(g - rg) / max(np.linalg.norm(rg), 1e-30)
            worst = max(worst, e)
            assert e <= 3e-2, (li, "weights", e)
            # biases: each channel a sum over the layer's voxels, relative L2
            # as for the weights plus the rounding scale of those sums (3e-3 of
            # the vector of sum |dy|), which dominates where a sum cancels (the
            # output layer's: its terms add up to the crop-face terms only)
            g, rg = grads[off:off + cout], ref_grads[off:off + cout]
            off += cout
            err = np.linalg.norm(g - rg)
            bound = 3e-2 * np.linalg.norm(rg) + 3e-3 * np.linalg.norm(abs_sums[li])
            worst_bias = max(worst_bias, float(err / bound))
            assert err <= bound, (li, "bias", err, np.linalg.norm(rg), np.linalg.norm(abs_sums[li]))
        print(f"PARITY train64 loss_rel={rel_loss:.3e} worst_layer_weight_grad_rel={worst:.3e} "
              f"worst_bias_error_over_bound={worst_bias:.3e}", flush=True)
        assert rel_loss <= 1e-3
        first = None
        for _ in range(20):
            assert L.smoke_nn_preconditioner_train_step(net, d_r, d_f, count, cd, 1e-3, C.byref(loss), None, None) == 0
            first = loss.value if first is None else first
        print(f"64^3 crops: loss {first:.4f} -> {loss.value:.4f} over 20 Adam steps", flush=True)
        assert loss.value < 0.7 * first, (first, loss.value)
    finally:
        hip.free(d_r)
        hip.free(d_f)
        L.smoke_nn_preconditioner_destroy(net)
