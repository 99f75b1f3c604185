"""The allocation-free network application
(smoke_nn_preconditioner_apply_static) against the allocating one
(smoke_nn_preconditioner_apply): same kernels on a static device workspace,
so the same bits -- on two calls in a row, the second on a smaller volume, and
again on the first size after that (nothing of an earlier call leaks into a
later one).  A volume beyond the workspace's limit is refused on the host."""
import numpy as np
import pytest

import bench
import smoke_abi as abi
from multilevel_reference import flags_of

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    abi.load_library()
    h = abi.Hip()
    if h.device_count() == 0:
        pytest.skip("no GPU")
    return h


def _volume(dims, seed):
    rng = np.random.default_rng(seed)
    W, H, D = dims
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    solid = (x - W * 0.6) ** 2 + (y - H * 0.4) ** 2 + (z - D * 0.5) ** 2 <= (0.2 * min(dims)) ** 2
    flags = flags_of(solid, int(np.flatnonzero(~solid.reshape(-1))[0]))
    r = rng.normal(size=(D, H, W)).astype(np.float32)
    r[flags == 0x40] = 0.0
    return r, flags


def _both(hip, net, dims, seed):
    r, flags = _volume(dims, seed)
    d_r, d_f = hip.upload(r), hip.upload(flags)
    d_a, d_b = hip.malloc(r.nbytes), hip.malloc(r.nbytes)
    try:
        net.apply(d_r, d_f, dims, d_a)
        rc = net.apply_static(d_r, d_f, dims, d_b)
        hip.synchronize()
        assert rc == 0
        a = hip.download(d_a, r.shape)
        b = hip.download(d_b, r.shape)
    finally:
        for p in (d_r, d_f, d_a, d_b):
            hip.free(p)
    return a, b, flags


def test_static_workspace_apply_matches_the_allocating_apply(hip):
    with abi.LearnedPreconditioner(seed=0, path=bench.WEIGHTS) as net:
        outs = {}
        for dims, seed in (((24, 20, 16), 1), ((12, 12, 8), 2), ((24, 20, 16), 3)):
            print(f"no-alloc apply: dims={dims} seed={seed}", flush=True)
            a, b, flags = _both(hip, net, dims, seed)
            assert np.isfinite(a).all() and np.abs(a).max() > 0
            np.testing.assert_array_equal(b, a, err_msg=str(dims))
            assert np.all(b[flags == 0x40] == 0.0)
            outs[(dims, seed)] = b
        # different inputs gave different outputs (the workspace is not stale)
        assert not np.array_equal(outs[((24, 20, 16), 1)], outs[((24, 20, 16), 3)])


def test_static_workspace_refuses_a_volume_beyond_its_limit(hip):
    with abi.LearnedPreconditioner(seed=0) as net:
        # 45^3 needs about 17.4 MB of the 16 MiB workspace: refused before any launch
        rc = net.apply_static(1, 1, (45, 45, 45), 1)
        assert rc == 2
        rc = net.apply_static(1, 1, (3, 8, 8), 1)   # below the network's minimum size
        assert rc == 2
