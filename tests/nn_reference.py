"""Float64 numpy restatement of the 3x3x3 convolutions of
visual-simulation-of-smoke-3d-nn.h (modes SAME, DOWN, UP; channels-last
volumes (B, D, H, W, C); weights [27][cin][cout], t = (tz * 3 + ty) * 3 + tx).
Test infrastructure only: the checker of the HIP kernels.

The data gradient is written as the adjoint of the forward gather
(scatter-add), not as the dual convolution the kernels use, so a wrong tap
flip or a DOWN/UP mix-up in the kernels shows up against it."""
from __future__ import annotations

import numpy as np

SAME, DOWN, UP = 0, 1, 2


def out_size(mode: int, n: int) -> int:
    return n if mode == SAME else ((n + 1) // 2 if mode == DOWN else 2 * n)


def source(mode: int, o: np.ndarray, t: int, n_in: int):
    """Input index and validity for output coordinates o through tap t."""
    if mode == SAME:
        i = o + t - 1
        ok = np.ones_like(o, bool)
    elif mode == DOWN:
        i = 2 * o + t - 1
        ok = np.ones_like(o, bool)
    else:
        q = o + 1 - t
        ok = (q % 2) == 0
        i = q // 2
    ok &= (i >= 0) & (i < n_in)
    return np.where(ok, i, 0), ok


def _taps(x_shape, out_dims, mode):
    _, D, H, W, _ = x_shape
    Do, Ho, Wo = out_dims
    for t in range(27):
        tx, ty, tz = t % 3, (t // 3) % 3, t // 9
        ix, vx = source(mode, np.arange(Wo), tx, W)
        iy, vy = source(mode, np.arange(Ho), ty, H)
        iz, vz = source(mode, np.arange(Do), tz, D)
        valid = vz[:, None, None] & vy[None, :, None] & vx[None, None, :]
        yield t, (iz, iy, ix), valid


def gather(x, idx, valid):
    iz, iy, ix = idx
    g = x[:, iz][:, :, iy][:, :, :, ix]
    return g * valid[None, :, :, :, None]


def forward(x, w, bias, mode, out_dims, skip=None, relu=False):
    B = x.shape[0]
    y = np.zeros((B, *out_dims, w.shape[2]))
    for t, idx, valid in _taps(x.shape, out_dims, mode):
        y += gather(x, idx, valid) @ w[t]
    if bias is not None:
        y += bias
    if skip is not None:
        y += skip
    return np.maximum(y, 0.0) if relu else y


def backward_data(dy, w, mode, in_shape):
    """Adjoint of forward(x) -> y with respect to x (scatter-add)."""
    dx = np.zeros(in_shape)
    out_dims = dy.shape[1:4]
    for t, (iz, iy, ix), valid in _taps(in_shape, out_dims, mode):
        contrib = (dy @ w[t].T) * valid[None, :, :, :, None]
        Z, Y, X = np.meshgrid(iz, iy, ix, indexing="ij")
        for b in range(dy.shape[0]):
            np.add.at(dx[b], (Z, Y, X), contrib[b])
    return dx


def backward_weights(x, dy, mode):
    out_dims = dy.shape[1:4]
    dw = np.zeros((27, x.shape[4], dy.shape[4]))
    for t, idx, valid in _taps(x.shape, out_dims, mode):
        g = gather(x, idx, valid)
        dw[t] = np.einsum("bzyxc,bzyxn->cn", g, dy)
    return dw, dy.sum(axis=(0, 1, 2, 3))
