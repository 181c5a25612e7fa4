"""gk_cons_sums_noalloc (the allocation-free conservation sums) against gk_cons_sums.

The caller owns the device memory: the test allocates the state and the output with the HIP
runtime the library itself uses (libamdhip64 from the ROCm install, through ctypes), copies a
state there, and compares the entry point's five sums with gk_cons_sums of a solver holding the
same state, bit for bit (same kernels, same summation order).  Two calls in a row on one GPU,
the second on a smaller state, in FP64 and FP32: the shared scratch array keeps the first call's
partials, which the second must overwrite completely.
"""
import ctypes as C

import numpy as np
import pytest

from gaukuk_amd import problems as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from gaukuk_amd import hip as H
    H.lib()   # loads the library and, with it, its HIP runtime
    return H


@pytest.fixture(scope="module")
def rt(hip):
    R = C.CDLL("libamdhip64.so.7")
    R.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    R.hipMalloc.restype = C.c_int
    R.hipFree.argtypes = [C.c_void_p]
    R.hipFree.restype = C.c_int
    R.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    R.hipMemcpy.restype = C.c_int
    R.hipDeviceSynchronize.restype = C.c_int
    return R


H2D, D2H = 1, 2


def _device_copy(rt, a):
    p = C.c_void_p()
    assert rt.hipMalloc(C.byref(p), a.nbytes) == 0
    assert rt.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
    return p


def _noalloc_sums(hip, rt, state, precision):
    """Sums of `state` ([5][ny][nx]) through gk_cons_sums_noalloc, on device copies."""
    a = np.ascontiguousarray(state, np.float64 if precision == 64 else np.float32)
    dstate = _device_copy(rt, a)
    out = np.full(5, np.nan)   # garbage in the output: the call must write all five
    dout = _device_copy(rt, out)
    try:
        st = hip.lib().gk_cons_sums_noalloc(dstate, a.shape[2], a.shape[1], 1, precision, dout, None)
        assert st == 0, hip.lib().gk_last_error(None)
        assert rt.hipDeviceSynchronize() == 0
        assert rt.hipMemcpy(out.ctypes.data, dout, out.nbytes, D2H) == 0
    finally:
        rt.hipFree(dstate)
        rt.hipFree(dout)
    return out


@pytest.mark.parametrize("precision", [64, 32])
def test_noalloc_sums_equal_cons_sums_two_calls(hip, rt, precision):
    cases = [P.kelvin_helmholtz(160, ny=120, seed=3, integrator=3, stepmax=15),
             P.blast(40, 26, p_ratio=30.0, radius=0.2, bc=(3, 0, 2, 1), stepmax=9)]   # smaller second call
    for prob in cases:
        s = hip.Solver(prob, precision=precision).run_to_end()
        state = s.download()
        ref = s.sums()
        got = _noalloc_sums(hip, rt, state, precision)
        print(f"FP{precision} {prob.nx}x{prob.ny}: gk_cons_sums {ref} noalloc {got}", flush=True)
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), (got, ref)
        s.close()


def test_noalloc_sums_reject_bad_arguments(hip):
    L = hip.lib()
    out = C.c_void_p(8)   # never dereferenced: the arguments are rejected first
    assert L.gk_cons_sums_noalloc(None, 4, 4, 1, 64, out, None) == 1
    assert L.gk_cons_sums_noalloc(out, 0, 4, 1, 64, out, None) == 1
    assert L.gk_cons_sums_noalloc(out, 4, 4, 1, 16, out, None) == 1
    assert b"precision" in L.gk_last_error(None)
