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


# ---------------------------------------------------------------- the preconditioner network
# (mode, cin, cout, input level, output level, relu), in parameter order
LAYERS = [
    (SAME, 8, 16, 0, 0, True), (SAME, 16, 16, 0, 0, True), (DOWN, 16, 32, 0, 1, True), (SAME, 32, 32, 1, 1, True),
    (DOWN, 32, 64, 1, 2, True), (SAME, 64, 64, 2, 2, True), (UP, 64, 32, 2, 1, True), (SAME, 32, 32, 1, 1, True),
    (UP, 32, 16, 1, 0, True), (SAME, 16, 16, 0, 0, True), (SAME, 16, 16, 0, 0, False),
]
IDENTITY = 0x40


def bf16(a):
    """Round to bfloat16 (nearest even) and back, in float64."""
    u = np.ascontiguousarray(np.asarray(a, np.float32)).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    return (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def split_parameters(flat):
    out, off = [], 0
    for _, cin, cout, *_ in LAYERS:
        w = flat[off:off + 27 * cin * cout].reshape(27, cin, cout)
        off += 27 * cin * cout
        b = flat[off:off + cout]
        off += cout
        out.append((w.astype(np.float64), b.astype(np.float64)))
    return out


def apply_A(v, flags):
    """Pressure operator on volumes (B, D, H, W): fluid rows only, identity
    rows and neighbours outside the volume read as zero."""
    ident = (flags & IDENTITY) != 0
    vv = np.where(ident, 0.0, v)
    bits = flags & 0x3F
    d = np.maximum(np.unpackbits(bits[..., None].astype(np.uint8), axis=-1).sum(-1), 1).astype(np.float64)
    out = d * vv
    pad = np.pad(vv, ((0, 0), (1, 1), (1, 1), (1, 1)))
    B, D, H, W = v.shape
    shifts = [(0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)]
    for bit, (dz, dy, dx) in enumerate(shifts):
        nb = pad[:, 1 + dz:D + 1 + dz, 1 + dy:H + 1 + dy, 1 + dx:W + 1 + dx]
        out -= np.where((bits >> bit) & 1, nb, 0.0)
    return np.where(ident, 0.0, out)


def unet_loss_and_gradients(flat, r, flags, rounding=True, loss="residual", abs_sums=None):
    """Loss mean_b ||A z - r||^2 / ||r||^2 (loss="residual") or
    mean_b (z^T A z / 2 - z^T r) / ||r||^2 over the fluid rows (loss="energy")
    and the parameter gradients of the preconditioner network, rounding every
    stored tensor to bfloat16 where the GPU does (rounding=False: exact
    float64 throughout, for finite-difference checks).  r, flags: (B, D, H, W).
    With a list `abs_sums`, appends per layer (in layer order) the sum over
    voxels of |dy| per output channel: the scale of the rounding error of any
    summation of that layer's bias gradient (which cancels where the terms
    nearly sum to zero, as the output layer's does on a large crop: its terms
    sum to (A 1)^T q, non-zero only next to the crop's faces and solids)."""
    bf16_ = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    f32 = (lambda a: np.asarray(a, np.float32).astype(np.float64)) if rounding else (lambda a: np.asarray(a, np.float64))
    P = split_parameters(flat)
    Wb = [(bf16_(w), f32(b)) for w, b in P]
    B, D, H, W = r.shape
    dims = [(D, H, W)]
    for _ in range(2):
        dims.append(tuple((n + 1) // 2 for n in dims[-1]))
    ident = (flags & IDENTITY) != 0
    s = np.sqrt((r.astype(np.float64) ** 2).reshape(B, -1).sum(1) / (D * H * W))
    s = np.where(s > 0, s, 1.0)
    h = bf16_((r / s[:, None, None, None]).astype(np.float32) if rounding else r / s[:, None, None, None])
    rn = np.where(ident, 0.0, h)
    x0 = np.zeros((B, D, H, W, 8))
    x0[..., 0] = rn
    x0[..., 1] = np.where(ident, 1.0, 0.0)

    def conv(li, x, skip=None):
        mode, cin, cout, li_in, li_out, relu = LAYERS[li]
        y = forward(x, Wb[li][0], Wb[li][1], mode, dims[li_out], skip=skip, relu=relu)
        return y

    T = {"in": x0}
    T["a0"] = bf16_(conv(0, T["in"]))
    T["e0"] = bf16_(conv(1, T["a0"]))
    T["d1"] = bf16_(conv(2, T["e0"]))
    T["e1"] = bf16_(conv(3, T["d1"]))
    T["d2"] = bf16_(conv(4, T["e1"]))
    T["e2"] = bf16_(conv(5, T["d2"]))
    T["up1"] = bf16_(conv(6, T["e2"], skip=T["e1"]))
    T["c1u"] = bf16_(conv(7, T["up1"]))
    T["up0"] = bf16_(conv(8, T["c1u"], skip=T["e0"]))
    T["c0u"] = bf16_(conv(9, T["up0"]))
    z = f32(conv(10, T["c0u"])[..., 0])
    q = np.where(ident, 0.0, apply_A(z, flags) - rn)
    rr = (rn ** 2).reshape(B, -1).sum(1)
    if loss == "energy":
        zf = np.where(ident, 0.0, z)
        ee = (0.5 * zf * apply_A(zf, flags) - zf * rn).reshape(B, -1).sum(1)
        loss = float(np.mean(ee / rr))
        g = (1.0 / B) * q / rr[:, None, None, None]
    else:
        qq = (q ** 2).reshape(B, -1).sum(1)
        loss = float(np.mean(qq / rr))
        g = (2.0 / B) * apply_A(q, flags) / rr[:, None, None, None]
    gout = np.zeros((B, D, H, W, 16))
    gout[..., 0] = bf16_(g.astype(np.float32) if rounding else g)

    grads = [None] * len(LAYERS)

    sums = [None] * len(LAYERS)

    def wgrad(li, x, dy):
        mode = LAYERS[li][0]
        grads[li] = backward_weights(x, dy, mode)
        sums[li] = np.abs(dy).reshape(-1, dy.shape[-1]).sum(0)

    def dgrad(li, dy, mask, add=None):
        mode = LAYERS[li][0]
        dx = backward_data(dy, Wb[li][0], mode, mask.shape)
        if add is not None:
            dx = dx + add
        return bf16_(dx * (mask > 0))

    wgrad(10, T["c0u"], gout)
    g0a = dgrad(10, gout, T["c0u"])
    wgrad(9, T["up0"], g0a)
    g0b = dgrad(9, g0a, T["up0"])
    wgrad(8, T["c1u"], g0b)
    g1a = dgrad(8, g0b, T["c1u"])
    wgrad(7, T["up1"], g1a)
    g1b = dgrad(7, g1a, T["up1"])
    wgrad(6, T["e2"], g1b)
    g2a = dgrad(6, g1b, T["e2"])
    wgrad(5, T["d2"], g2a)
    g2b = dgrad(5, g2a, T["d2"])
    wgrad(4, T["e1"], g2b)
    g1a = dgrad(4, g2b, T["e1"], add=g1b)
    wgrad(3, T["d1"], g1a)
    g1b = dgrad(3, g1a, T["d1"])
    wgrad(2, T["e0"], g1b)
    g0a = dgrad(2, g1b, T["e0"], add=g0b)
    wgrad(1, T["a0"], g0a)
    g0b = dgrad(1, g0a, T["a0"])
    wgrad(0, T["in"], g0b)
    flat_grads = np.concatenate([np.concatenate([dw.ravel(), db.ravel()]) for dw, db in grads])
    if abs_sums is not None:
        abs_sums.extend(sums)
    return loss, flat_grads
