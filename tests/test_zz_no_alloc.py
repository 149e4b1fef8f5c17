"""rp_mlp_infer_noalloc (the inference entry point that allocates nothing)
against rp_mlp_infer on the same parameters and rows, through the C ABI:
outputs equal to the bit (same packing, same kernel), on two calls in a row
-- the second with fewer rows, so a stale row or weight left in the shared
scratch array from the first call would show -- plus the size limit and
argument errors.  The test allocates its device buffers and stream with the
HIP runtime the library itself is linked to (the one loaded in this
process), so no second runtime enters the process; the entry point itself
allocates nothing."""
import ctypes as C

import numpy as np
import pytest

from rusty_particles_amd import abi, native

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _libs():
    L = native.lib()
    L.rp_mlp_infer.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_size_t, C.POINTER(C.c_float)]
    L.rp_mlp_infer.restype = C.c_int32
    L.rp_mlp_infer_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]
    L.rp_mlp_infer_noalloc.restype = C.c_int32
    L.rp_surrogate_last_error.restype = C.c_char_p
    path = None
    with open("/proc/self/maps") as fh:
        for line in fh:
            if "libamdhip64" in line:
                path = line.split()[-1]
                break
    assert path, "the library's HIP runtime is not loaded"
    H = C.CDLL(path)
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipFree.argtypes = [C.c_void_p]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    H.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    H.hipStreamSynchronize.argtypes = [C.c_void_p]
    H.hipStreamDestroy.argtypes = [C.c_void_p]
    for f in (H.hipMalloc, H.hipFree, H.hipMemcpy, H.hipStreamCreate, H.hipStreamSynchronize, H.hipStreamDestroy):
        f.restype = C.c_int
    return L, H


def _dev(H, nbytes):
    p = C.c_void_p()
    assert H.hipMalloc(C.byref(p), max(int(nbytes), 4)) == 0
    return p


def _host_infer(L, params, x):
    y = np.zeros((x.shape[0], 2), np.float32)
    fp = C.POINTER(C.c_float)
    assert L.rp_mlp_infer(params.ctypes.data_as(fp), x.ctypes.data_as(fp), x.shape[0], y.ctypes.data_as(fp)) == 0
    return y


def test_noalloc_matches_allocating_inference():
    L, H = _libs()
    assert native.device_count() > 0
    rng = np.random.default_rng(11)
    stream = C.c_void_p()
    assert H.hipStreamCreate(C.byref(stream)) == 0
    bufs = []
    try:
        for call, (rows, seed) in enumerate(((1000, 1), (300, 2))):
            print(f"case call={call} rows={rows}", flush=True)
            params = np.random.default_rng(seed).normal(0.0, 0.3, abi.RP_MLP_PARAMS).astype(np.float32)
            x = rng.uniform(-1.0, 1.0, (rows, 4)).astype(np.float32)
            want = _host_infer(L, params, x)
            p_dev, x_dev, y_dev = _dev(H, params.nbytes), _dev(H, x.nbytes), _dev(H, rows * 2 * 4)
            bufs += [p_dev, x_dev, y_dev]
            nan = np.full((rows, 2), np.nan, np.float32)  # every row must be written by this call
            assert H.hipMemcpy(p_dev, params.ctypes.data, params.nbytes, H2D) == 0
            assert H.hipMemcpy(x_dev, x.ctypes.data, x.nbytes, H2D) == 0
            assert H.hipMemcpy(y_dev, nan.ctypes.data, nan.nbytes, H2D) == 0
            code = L.rp_mlp_infer_noalloc(p_dev, x_dev, rows, y_dev, stream)
            assert code == 0, L.rp_surrogate_last_error()
            assert H.hipStreamSynchronize(stream) == 0
            got = np.zeros((rows, 2), np.float32)
            assert H.hipMemcpy(got.ctypes.data, y_dev, got.nbytes, D2H) == 0
            assert np.array_equal(got, want), (call, np.abs(got - want).max())
    finally:
        for b in bufs:
            H.hipFree(b)
        H.hipStreamDestroy(stream)


def test_noalloc_limits_and_errors():
    L, H = _libs()
    p_dev, x_dev, y_dev = _dev(H, abi.RP_MLP_PARAMS * 4), _dev(H, 16 * 16), _dev(H, 16 * 8)
    try:
        too_many = (1 << 28) + 1  # RP_MLP_NOALLOC_MAX_ROWS + 1: refused before any launch
        assert L.rp_mlp_infer_noalloc(p_dev, x_dev, too_many, y_dev, None) == abi.RP_ERR_INVALID
        assert b"RP_MLP_NOALLOC_MAX_ROWS" in L.rp_surrogate_last_error()
        assert L.rp_mlp_infer_noalloc(None, x_dev, 16, y_dev, None) == abi.RP_ERR_INVALID
        assert L.rp_mlp_infer_noalloc(p_dev, None, 16, y_dev, None) == abi.RP_ERR_INVALID
        assert L.rp_mlp_infer_noalloc(p_dev, None, 0, None, None) == 0  # zero rows: nothing to do
    finally:
        for b in (p_dev, x_dev, y_dev):
            H.hipFree(b)
