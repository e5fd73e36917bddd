"""fednl_newton_solve_noalloc: the Newton solve on device pointers with its
workspace in the library's own device array (no allocation), against the
allocating fednl_newton_solve on the same inputs (same kernels, so the same
bits). Two calls in a row, the second with a smaller d, on the null stream
and on a stream of the caller's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _spd_packed(rng, d):
    A = rng.standard_normal((d, d))
    Hs = A @ A.T / d + 0.1 * np.eye(d)
    return np.array([Hs[i, j] for j in range(d) for i in range(j + 1)])


def _dev(torch, a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)


def _solve_noalloc(torch, E, H_up, d, shift, b, stream=None):
    H_t = _dev(torch, H_up, torch.float64)
    b_t = _dev(torch, b, torch.float64)
    dx_t = torch.full((d,), float("nan"), dtype=torch.float64, device="cuda")
    ok_t = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    handle = None if stream is None else stream.cuda_stream
    E.newton_solve_noalloc(d, H_t.data_ptr(), shift, b_t.data_ptr(), dx_t.data_ptr(), ok_t.data_ptr(), handle)
    torch.cuda.synchronize()
    return int(ok_t.item()), dx_t.cpu().numpy()


@pytest.mark.parametrize("sizes", [(300, 100), (512, 37), (64, 1)])
def test_noalloc_matches_allocating_solve(sizes):
    import torch
    from ufednl_amd import engine as E
    rng = np.random.default_rng(sum(sizes))
    stream = torch.cuda.Stream()  # default priority
    for q, d in enumerate(sizes):
        H_up = _spd_packed(rng, d)
        b = rng.standard_normal(d)
        print(f"noalloc newton d={d}", flush=True)
        ok_ref, dx_ref = E.newton_solve(H_up, d, 1e-3, b)
        ok, dx = _solve_noalloc(torch, E, H_up, d, 1e-3, b, stream if q == 1 else None)
        assert ok_ref and ok == 1
        np.testing.assert_array_equal(dx, dx_ref)


def test_noalloc_reports_failure_then_recovers():
    import torch
    from ufednl_amd import engine as E
    d = 6
    Hs = -np.eye(d)
    H_up = np.array([Hs[i, j] for j in range(d) for i in range(j + 1)])
    ok, dx = _solve_noalloc(torch, E, H_up, d, 0.0, np.ones(d))
    assert ok == 0
    assert np.all(dx == 0.0)
    # the next call starts from a clean status record
    rng = np.random.default_rng(5)
    H_up = _spd_packed(rng, 40)
    b = rng.standard_normal(40)
    ok_ref, dx_ref = E.newton_solve(H_up, 40, 1e-3, b)
    ok, dx = _solve_noalloc(torch, E, H_up, 40, 1e-3, b)
    assert ok == 1
    np.testing.assert_array_equal(dx, dx_ref)


def test_noalloc_rejects_oversize():
    from ufednl_amd import engine as E
    with pytest.raises(E.FednlError, match="d > 512"):
        E.newton_solve_noalloc(513, 1, 1e-3, 1, 1, 1, None)
