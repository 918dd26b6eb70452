"""hz_mlp_forward_noalloc (device pointers, static scratch, no allocation)
against hz_mlp_forward, the allocating entry point it is a variant of: the
same kernels, so the outputs must agree bit for bit, on two calls in a row
with the second one smaller, and a batch above the stated maximum is refused."""
import ctypes as C

import numpy as np
import pytest

from hazmath_amd import capi

pytestmark = pytest.mark.gpu


def _hip():
    from test_gpu_parity import _hip_runtime
    return _hip_runtime()


def _upload(hip, arr):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), C.c_size_t(max(arr.nbytes, 4))) == 0
    assert hip.hipMemcpy(p, arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes), 1) == 0
    return p


def test_mlp_forward_noalloc_matches_allocating_entry(lib):
    hip = _hip()
    theta = np.zeros(capi.MLP_PARAMS, np.float32)
    lib.hz_mlp_init(capi.fptr(theta), 21)
    theta[1024:1088] = np.linspace(-0.2, 0.2, 64, dtype=np.float32)   # nonzero first-layer biases
    rng = np.random.default_rng(5)
    cap = 5000
    dth = _upload(hip, theta)
    dx = _upload(hip, np.zeros((cap, capi.MLP_INPUTS), np.float32))
    dy = _upload(hip, np.zeros((cap, capi.MLP_OUTPUTS), np.float32))
    try:
        for B in (cap, 301):
            print(f"[no-alloc] call with B={B}", flush=True)
            x = rng.standard_normal((B, capi.MLP_INPUTS)).astype(np.float32)
            ref = np.zeros((B, capi.MLP_OUTPUTS), np.float32)
            assert lib.hz_mlp_forward(capi.fptr(theta), capi.fptr(x), B, capi.fptr(ref)) == 0
            assert hip.hipMemcpy(dx, x.ctypes.data_as(C.c_void_p), C.c_size_t(x.nbytes), 1) == 0
            assert lib.hz_mlp_forward_noalloc(dth, dx, B, dy, None) == 0
            assert hip.hipDeviceSynchronize() == 0
            out = np.zeros((B, capi.MLP_OUTPUTS), np.float32)
            assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), dy, C.c_size_t(out.nbytes), 2) == 0
            assert np.array_equal(out, ref), np.abs(out - ref).max()
        too_many = capi.MLP_NOALLOC_MAX_ROWS + 1
        assert lib.hz_mlp_forward_noalloc(dth, dx, too_many, dy, None) == capi.ERROR_INPUT_PAR
        assert lib.hz_mlp_forward_noalloc(dth, dx, -1, dy, None) == capi.ERROR_INPUT_PAR
    finally:
        for p in (dth, dx, dy):
            hip.hipFree(p)
