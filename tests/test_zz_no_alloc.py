"""The allocation-free forward pass of the warm-start network (GPU).

erosion_mlp_forward_noalloc takes device pointers and a stream and keeps its
hidden activations in a file-scope device array; it must return what
erosion_mlp_forward (host pointers, library-allocated buffers) returns, bit
for bit (same kernels, same summation order), on two calls in a row, the
second with fewer samples, and refuse a network wider than its scratch.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
SHAPE = (128, 256, 3, 69)


def _noalloc():
    from erosion_amd import stokes
    L = stokes.lib()
    f = L.erosion_mlp_forward_noalloc
    f.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p]
    f.restype = ctypes.c_int
    L.erosion_noalloc_last_error.restype = ctypes.c_char_p
    return f, L


class _Hip:
    """Device buffers for the test, straight from the HIP runtime."""

    def __init__(self):
        from erosion_amd import stokes
        stokes.lib()  # loads the HIP runtime with the library
        h = ctypes.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        for fn in (h.hipMalloc, h.hipMemcpy, h.hipFree, h.hipDeviceSynchronize,
                   h.hipStreamCreate, h.hipStreamDestroy, h.hipStreamSynchronize):
            fn.restype = ctypes.c_int
        self.h = h

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, shape, dtype=np.float32):
        out = np.empty(shape, dtype=dtype)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out


def _device_forward(params, x, shape):
    f, L = _noalloc()
    hip = _Hip()
    dp, dx = hip.upload(params), hip.upload(x)
    dy = hip.upload(np.full((x.shape[0], shape[3]), np.nan, dtype=np.float32))
    stream = ctypes.c_void_p()
    assert hip.h.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        rc = f(*shape, dp, x.shape[0], dx, dy, stream)
        assert hip.h.hipStreamSynchronize(stream) == 0
        assert rc == 0, L.erosion_noalloc_last_error().decode()
        return hip.download(dy, (x.shape[0], shape[3]))
    finally:
        hip.h.hipStreamDestroy(stream)
        for p in (dp, dx, dy):
            hip.h.hipFree(p)


def test_noalloc_forward_matches_allocating_forward():
    from erosion_amd import learned as LR
    rng = np.random.default_rng(21)
    params = (LR.init_params(SHAPE, seed=5) + 0.01 * rng.standard_normal(LR.param_count(SHAPE))
              ).astype(np.float32)
    # 9000 samples: three chunks of the scratch (4096 rows of 256 units), then 300
    for ns in (9000, 300):
        x = rng.standard_normal((ns, SHAPE[0])).astype(np.float32)
        y_ref = LR.mlp_forward(params, x, SHAPE)
        y = _device_forward(params, x, SHAPE)
        assert np.all(np.isfinite(y))
        assert np.array_equal(y, y_ref), np.max(np.abs(y - y_ref))


def test_noalloc_forward_refuses_a_network_wider_than_its_scratch():
    from erosion_amd import learned as LR
    f, L = _noalloc()
    hip = _Hip()
    shape = (16, 40000, 1, 4)  # hidden width above the 32768 units the scratch allows
    # buffers of the full size of this network, so that nothing reads past them
    bufs = [hip.upload(np.zeros(k, dtype=np.float32))
            for k in (LR.param_count(shape), shape[0], shape[3])]
    try:
        rc = f(*shape, bufs[0], 1, bufs[1], bufs[2], None)
        assert rc == -1
        assert "exceeds the forward scratch" in L.erosion_noalloc_last_error().decode()
    finally:
        for p in bufs:
            hip.h.hipFree(p)
