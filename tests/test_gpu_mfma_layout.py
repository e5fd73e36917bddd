"""Pins the lane layout of v_mfma_f64_16x16x4f64 that k_hessian_f64 uses:
A[m][k] from lane 16k+m, B[k][n] from lane 16k+n, D[m][n] in lane 16(m%4)+n, item m//4."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_mfma_f64_layout():
    from ufednl_amd.engine import lib
    L = lib()
    L.fednl_probe_mfma_f64.argtypes = [ctypes.POINTER(ctypes.c_double)]
    out = np.zeros(768)
    assert L.fednl_probe_mfma_f64(out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    r1, r2, r3 = out[:256].reshape(64, 4), out[256:512].reshape(64, 4), out[512:].reshape(64, 4)
    print("r1", r1.tolist())
    print("r2", r2.tolist())
    print("r3", r3.tolist())
    for lane in range(64):
        for r in range(4):
            m = 4 * r + lane // 16
            n = lane % 16
            # r1: sum over k of A[m][k] = (16k + m + 1)
            assert r1[lane, r] == sum(16 * kk + m + 1 for kk in range(4)), (lane, r, r1[lane, r])
            assert r2[lane, r] == sum(16 * kk + n + 1 for kk in range(4)), (lane, r, r2[lane, r])
            assert r3[lane, r] == m + 1, (lane, r, r3[lane, r])
