"""Pins the float64 convolution restatement (nn_reference.py) that checks the
matrix-core kernels: brute-force loops on tiny volumes for every mode, and
the adjoint identities <dy, conv(x)> = <conv^T(dy), x> = <dW(x, dy), W> that
tie its data and weight gradients to its forward pass."""
import numpy as np
import pytest

import nn_reference as ref


def brute(x, w, mode, out_dims):
    B, D, H, W, _ = x.shape
    Do, Ho, Wo = out_dims
    y = np.zeros((B, Do, Ho, Wo, w.shape[2]))
    for b in range(B):
        for oz in range(Do):
            for oy in range(Ho):
                for ox in range(Wo):
                    for t in range(27):
                        tx, ty, tz = t % 3, (t // 3) % 3, t // 9
                        src = []
                        for o, tt, n in ((ox, tx, W), (oy, ty, H), (oz, tz, D)):
                            if mode == ref.SAME:
                                i = o + tt - 1
                            elif mode == ref.DOWN:
                                i = 2 * o + tt - 1
                            else:
                                q = o + 1 - tt
                                i = q // 2 if q % 2 == 0 else -1
                            src.append(i if 0 <= i < n else None)
                        if None in src:
                            continue
                        ix, iy, iz = src
                        y[b, oz, oy, ox] += x[b, iz, iy, ix] @ w[t]
    return y


@pytest.mark.parametrize("mode,size", [(ref.SAME, (4, 3, 2)), (ref.DOWN, (5, 4, 3)), (ref.UP, (3, 2, 2))])
def test_forward_matches_brute_force_and_adjoints_hold(mode, size):
    rng = np.random.default_rng(mode)
    W, H, D = size
    out = tuple(ref.out_size(mode, s) for s in size)[::-1]
    x = rng.normal(size=(2, D, H, W, 3))
    w = rng.normal(size=(27, 3, 4))
    y = ref.forward(x, w, None, mode, out)
    np.testing.assert_allclose(y, brute(x, w, mode, out), rtol=1e-12, atol=1e-12)
    dy = rng.normal(size=y.shape)
    dx = ref.backward_data(dy, w, mode, x.shape)
    dw, db = ref.backward_weights(x, dy, mode)
    inner = float((dy * y).sum())
    assert abs(inner - float((dx * x).sum())) <= 1e-10 * abs(inner)
    assert abs(inner - float((dw * w).sum())) <= 1e-10 * abs(inner)
    np.testing.assert_allclose(db, dy.sum(axis=(0, 1, 2, 3)))


@pytest.mark.parametrize("loss_kind", ["residual", "energy"])
def test_network_gradients_match_finite_differences(loss_kind):
    """The float64 network restatement (rounding off) against central
    differences of its own loss, for parameters of every layer."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).parent))
    from test_gpu_preconditioner import volumes
    r, flags = volumes(1, (8, 6, 4), 5)
    rng = np.random.default_rng(1)
    sizes = [27 * ci * co + co for _, ci, co, *_ in ref.LAYERS]
    params = np.zeros(sum(sizes))
    off = 0
    for (_, ci, co, *_), n in zip(ref.LAYERS, sizes):
        params[off:off + n] = rng.normal(size=n) * np.sqrt(2.0 / (27 * ci))
        off += n
    loss, grads = ref.unet_loss_and_gradients(params, r, flags, rounding=False, loss=loss_kind)
    off = 0
    for li, n in enumerate(sizes):
        # the largest-gradient weight of the layer and its first bias
        w_part = grads[off:off + n - ref.LAYERS[li][2]]
        picks = [off + int(np.argmax(np.abs(w_part))), off + n - ref.LAYERS[li][2]]
        for k in picks:
            h = 1e-6 * max(1.0, abs(params[k]))
            p1, p2 = params.copy(), params.copy()
            p1[k] += h
            p2[k] -= h
            fd = (ref.unet_loss_and_gradients(p1, r, flags, rounding=False, loss=loss_kind)[0] -
                  ref.unet_loss_and_gradients(p2, r, flags, rounding=False, loss=loss_kind)[0]) / (2 * h)
            assert abs(fd - grads[k]) <= 1e-4 * max(abs(fd), 1e-3 * abs(loss)), (li, k, fd, grads[k])
        off += n


def test_energy_loss_is_minimised_by_the_exact_solve():
    """The energy loss of z = A^-1 r is -r^T A^-1 r / 2 / ||r||^2 and any other
    z scores higher by ||z - A^-1 r||_A^2 / 2 / ||r||^2 (closed form of the
    quadratic), checked on the operator restatement with a dense solve."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).parent))
    from test_gpu_preconditioner import volumes
    r, flags = volumes(1, (5, 4, 3), 2)
    r = r.astype(np.float64)
    fluid = (flags[0] & ref.IDENTITY) == 0
    n = r[0].size
    A = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        A[:, k] = ref.apply_A(e.reshape(1, *r.shape[1:]), flags).ravel()
    idx = np.flatnonzero(fluid.ravel())
    Af = A[np.ix_(idx, idx)]
    np.testing.assert_allclose(Af, Af.T)
    rf = r.ravel()[idx]
    zs = np.linalg.solve(Af, rf)

    def energy(zf):
        return 0.5 * zf @ Af @ zf - zf @ rf

    rng = np.random.default_rng(3)
    d = rng.normal(size=zs.shape)
    assert abs(energy(zs) + 0.5 * rf @ zs) <= 1e-12 * abs(rf @ zs)
    assert abs(energy(zs + d) - energy(zs) - 0.5 * d @ Af @ d) <= 1e-10 * abs(d @ Af @ d)
