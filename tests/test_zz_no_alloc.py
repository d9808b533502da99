"""galp_mlp_loss_grad_noalloc (the allocation-free variant of
galp_mlp_loss_grad, for callers that cannot allocate device memory) against
galp_mlp_loss_grad on the same data: two calls in a row on device-resident
inputs, the second with fewer samples, so the second reuses the scratch the
first left behind; on the null stream and on a stream of the caller's; and
the size limit is refused with a clear error.  The test's device buffers come
from the HIP runtime libgalp itself is linked against (ctypes)."""
import ctypes

import numpy as np
import pytest

from galp_amd import _abi

pytestmark = pytest.mark.gpu


def hip():
    _abi.lib()   # libgalp first: the HIP runtime it links is the one to use
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return h


H2D, D2H = 1, 2


class DevArray:
    def __init__(self, h, arr):
        self.h, self.nbytes = h, arr.nbytes
        self.p = ctypes.c_void_p()
        assert h.hipMalloc(ctypes.byref(self.p), max(self.nbytes, 4)) == 0
        assert h.hipMemcpy(self.p, arr.ctypes.data, self.nbytes, H2D) == 0

    def get(self, dtype, count):
        out = np.empty(count, dtype)
        assert self.h.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.h.hipFree(self.p)


def rand_x(n, seed):
    rng = np.random.default_rng(seed)
    x = np.empty((n, 4), np.float32)
    x[:, 0] = rng.uniform(6, 20, n)
    x[:, 1] = rng.uniform(0, 12, n)
    x[:, 2] = x[:, 1] + rng.uniform(0, 8, n)
    x[:, 3] = np.where(rng.random(n) < 0.2, -1.0, rng.uniform(0, 14, n))
    return x


@pytest.mark.parametrize("hidden,own_stream", [(32, False), (8, True)])
def test_noalloc_matches_allocating_entry_point(ctx, hidden, own_stream):
    print(f"[noalloc] hidden={hidden} own_stream={own_stream}", flush=True)
    h = hip()
    stream = ctypes.c_void_p()
    if own_stream:
        assert h.hipStreamCreate(ctypes.byref(stream)) == 0
    p = _abi.mlp_init(hidden, 3 + hidden)
    for n in (32768, 1000):   # the second call is smaller: stale scratch beyond it must not matter
        x = rand_x(n, n)
        y = (0.4 * x[:, 0] + 0.2 * x[:, 3]).astype(np.float32)
        loss_ref, grad_ref = _abi.mlp_loss_grad(ctx, hidden, p, x, y)
        bufs = [DevArray(h, a) for a in (p, x, y, np.full(len(p), np.nan, np.float32),
                                         np.full(1, np.nan, np.float32))]
        dp, dx, dy, dg, dl = bufs
        rc = _abi.lib().galp_mlp_loss_grad_noalloc(ctx.h, hidden, dp.p, dx.p, dy.p, n, dg.p, dl.p, stream)
        assert rc == 0, _abi.lib().galp_last_error(ctx.h)
        assert h.hipStreamSynchronize(stream) == 0
        grad, loss = dg.get(np.float32, len(p)), float(dl.get(np.float32, 1)[0])
        for b in bufs:
            b.free()
        # same kernels, same block order: the same bits
        assert np.array_equal(grad, grad_ref), f"n={n}"
        assert loss == loss_ref, f"n={n}"
    if own_stream:
        h.hipStreamDestroy(stream)


def test_noalloc_refuses_beyond_its_scratch(ctx):
    h = hip()
    n = 32769
    d = DevArray(h, np.zeros(4 * n + 2000, np.float32))
    rc = _abi.lib().galp_mlp_loss_grad_noalloc(ctx.h, 32, d.p, d.p, d.p, n, d.p, d.p, None)
    d.free()
    assert rc == -1   # GALP_EINVAL
    assert b"32768" in _abi.lib().galp_last_error(ctx.h)
