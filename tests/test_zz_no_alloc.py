"""fednl_newton_solve_noalloc: the Newton solve on device pointers with its
workspace in the library's own device array (no allocation), against the
allocating fednl_newton_solve on the same inputs (same kernels, so the same
bits). Two calls in a row, the second with a smaller d, on the null stream
and on a stream of the caller's.

The caller's device buffers and stream come from the HIP runtime the library
itself links (loaded through ctypes), so the test adds no second runtime to
the process."""
import ctypes
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    from ufednl_amd import engine as E
    E.lib()  # the library, and with it its HIP runtime, is loaded first
    out = subprocess.run(["ldd", E.lib_path()], capture_output=True, text=True).stdout
    m = re.search(r"libamdhip64\.so\S*\s+=>\s+(\S+)", out)
    path = m.group(1) if m else "libamdhip64.so"
    h = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    h.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    h.hipFree.argtypes = [vp]
    h.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    h.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
    h.hipStreamDestroy.argtypes = [vp]
    h.hipStreamSynchronize.argtypes = [vp]
    return h


class _Dev:
    """A device buffer of the test (the caller's input or output)."""

    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), max(nbytes, 8)) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D) == 0

    def get(self, dtype, count):
        a = np.empty(count, dtype=dtype)
        assert self.hip.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def _spd_packed(rng, d):
    A = rng.standard_normal((d, d))
    Hs = A @ A.T / d + 0.1 * np.eye(d)
    return np.concatenate([Hs[: j + 1, j] for j in range(d)])


def _solve_noalloc(hip, E, H_up, d, shift, b, stream=None):
    bufs = [_Dev(hip, 8 * len(H_up)), _Dev(hip, 8 * d), _Dev(hip, 8 * d), _Dev(hip, 4)]
    Hd, bd, dxd, okd = bufs
    Hd.put(np.asarray(H_up, dtype=np.float64))
    bd.put(np.asarray(b, dtype=np.float64))
    dxd.put(np.full(d, np.nan))
    okd.put(np.array([-1], dtype=np.int32))
    E.newton_solve_noalloc(d, Hd.p.value, shift, bd.p.value, dxd.p.value, okd.p.value,
                           None if stream is None else stream.value)
    if stream is not None:
        assert hip.hipStreamSynchronize(stream) == 0
    else:
        assert hip.hipDeviceSynchronize() == 0
    ok = int(okd.get(np.int32, 1)[0])
    dx = dxd.get(np.float64, d)
    for q in bufs:
        q.free()
    return ok, dx


@pytest.mark.parametrize("sizes", [(300, 100), (512, 37), (64, 1)])
def test_noalloc_matches_allocating_solve(sizes):
    from ufednl_amd import engine as E
    hip = _hip()
    rng = np.random.default_rng(sum(sizes))
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        for q, d in enumerate(sizes):
            H_up = _spd_packed(rng, d)
            b = rng.standard_normal(d)
            print(f"noalloc newton d={d}", flush=True)
            ok_ref, dx_ref = E.newton_solve(H_up, d, 1e-3, b)
            ok, dx = _solve_noalloc(hip, E, H_up, d, 1e-3, b, stream if q == 1 else None)
            assert ok_ref and ok == 1
            np.testing.assert_array_equal(dx, dx_ref)
    finally:
        hip.hipStreamDestroy(stream)


def test_noalloc_reports_failure_then_recovers():
    from ufednl_amd import engine as E
    hip = _hip()
    d = 6
    Hs = -np.eye(d)
    H_up = np.concatenate([Hs[: j + 1, j] for j in range(d)])
    ok, dx = _solve_noalloc(hip, E, H_up, d, 0.0, np.ones(d))
    assert ok == 0
    assert np.all(dx == 0.0)
    # the next call starts from a clean status record
    rng = np.random.default_rng(5)
    H_up = _spd_packed(rng, 40)
    b = rng.standard_normal(40)
    ok_ref, dx_ref = E.newton_solve(H_up, 40, 1e-3, b)
    ok, dx = _solve_noalloc(hip, E, H_up, 40, 1e-3, b)
    assert ok == 1
    np.testing.assert_array_equal(dx, dx_ref)


def test_noalloc_rejects_oversize():
    from ufednl_amd import engine as E
    with pytest.raises(E.FednlError, match="d > 512"):
        E.newton_solve_noalloc(513, 1, 1e-3, 1, 1, 1, None)
