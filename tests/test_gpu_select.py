"""Index selection on the GPU vs the oracle on identical inputs (bit-exact).

Cases follow the reference's selectors: TopK / TopLEK heap scan
(SpecialMathRoutinesForMatrix.h:229-511), RandK partial shuffle (:95-107,
Shuffle.h:31-93), RandSeqK window (:137-169). Inputs include ties, exact
zeros of both signs and the k >= S edge.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _deltas(rng, kind, nc, S):
    if kind == "gauss":
        return rng.standard_normal((nc, S))
    if kind == "ties":
        return rng.integers(-3, 4, (nc, S)).astype(np.float64)
    if kind == "zeros":
        return np.where(rng.random((nc, S)) < 0.5, 0.0, -0.0)
    if kind == "round1":
        return np.round(rng.standard_normal((nc, S)), 1)
    raise ValueError(kind)


CASES = [(5, 3), (12, 40), (30, 240), (68, 544), (123, 984), (300, 2400)]


@pytest.mark.parametrize("d,k", CASES)
@pytest.mark.parametrize("kind", ["gauss", "ties", "zeros", "round1"])
def test_topk_bit_exact(oracle, d, k, kind):
    from ufednl_amd import engine as E
    rng = np.random.default_rng(d * 1000 + k + len(kind))
    S = d * (d + 1) // 2
    nc = 3
    delta = _deltas(rng, kind, nc, S)
    print(f"topk d={d} k={k} kind={kind}", flush=True)
    pos, cnt, _, _ = E.select("topk", delta, d, k)
    for c in range(nc):
        ref = oracle.topk_packed(delta[c], k)
        assert cnt[c] == len(ref)
        np.testing.assert_array_equal(pos[c, :cnt[c]], ref)


# past the LDS-resident heap (k > 8192) and past the LDS bitmap (S > 2^20):
# the heap in global scratch, the bitmap built in global memory
WIDE = [(1100, 9000), (1500, 3000), (1500, 12000), (2048, 16384)]


@pytest.mark.parametrize("d,k", WIDE)
@pytest.mark.parametrize("kind", ["gauss", "ties", "zeros"])
def test_topk_wide_bit_exact(oracle, d, k, kind):
    test_topk_bit_exact(oracle, d, k, kind)


@pytest.mark.parametrize("d,k", [(1100, 9000), (2048, 16384)])
@pytest.mark.parametrize("kind", ["gauss", "round1"])
@pytest.mark.parametrize("workers", [1, 2])
def test_toplek_wide_bit_exact(oracle, d, k, kind, workers):
    test_toplek_bit_exact(oracle, d, k, kind, workers)


@pytest.mark.parametrize("d,k", [(5, 3), (12, 40), (30, 100), (68, 544), (123, 984)])
@pytest.mark.parametrize("kind", ["gauss", "ties", "round1", "zeros"])
@pytest.mark.parametrize("workers", [1, 2])
def test_toplek_bit_exact(oracle, d, k, kind, workers):
    from ufednl_amd import engine as E
    rng = np.random.default_rng(7 * d + k + workers)
    S = d * (d + 1) // 2
    nc = 5
    delta = _deltas(rng, kind, nc, S)
    lek = np.full(workers, 1234, dtype=np.uint32)
    print(f"toplek d={d} k={k} kind={kind} W={workers}", flush=True)
    pos, cnt, _, lek_after = E.select("toplek", delta, d, k, toplek_rng=lek, workers=workers)
    streams = [1234] * workers
    for c in range(nc):
        ref, streams[c % workers] = oracle.toplek_packed(delta[c], d, k, k * (1.0 / S), streams[c % workers])
        assert cnt[c] == len(ref), (c, cnt[c], len(ref))
        np.testing.assert_array_equal(pos[c, :cnt[c]], ref)
    np.testing.assert_array_equal(lek_after, np.array(streams, dtype=np.uint32))


@pytest.mark.parametrize("d,k", [(5, 3), (5, 15), (12, 40), (68, 544), (123, 984), (300, 2400), (1500, 3000),
                                 (2048, 16384)])
def test_randk_bit_exact(oracle, d, k):
    from ufednl_amd import engine as E
    S = d * (d + 1) // 2
    nc = 4
    rng0 = np.array([123 + c for c in range(nc)], dtype=np.uint32)
    delta = np.zeros((nc, S))
    print(f"randk d={d} k={k}", flush=True)
    states = rng0.copy()
    for rnd in range(3):
        pos, cnt, states_gpu, _ = E.select("randk", delta, d, k, rng=states)
        for c in range(nc):
            ref, st = oracle.randk_pos(int(states[c]), S, k)
            np.testing.assert_array_equal(pos[c], ref)
            assert int(states_gpu[c]) == st
        states = states_gpu


@pytest.mark.parametrize("d,k", [(5, 3), (12, 78), (68, 544), (300, 2400), (2048, 16384)])
def test_randseqk_bit_exact(oracle, d, k):
    from ufednl_amd import engine as E
    S = d * (d + 1) // 2
    nc = 6
    states = np.array([123 + c for c in range(nc)], dtype=np.uint32)
    delta = np.zeros((nc, S))
    print(f"seqk d={d} k={k}", flush=True)
    for rnd in range(4):
        pos, cnt, states_gpu, _ = E.select("seqk", delta, d, k, rng=states)
        for c in range(nc):
            ref, st = oracle.randseqk_pos(int(states[c]), S, k)
            np.testing.assert_array_equal(pos[c], np.sort(ref))
            assert int(states_gpu[c]) == st
        states = states_gpu


def test_newton_solve_matches_oracle(oracle):
    from ufednl_amd import engine as E
    rng = np.random.default_rng(3)
    for d in (1, 7, 33, 64, 100, 300, 321, 480, 511, 512):
        A = rng.standard_normal((d, d))
        Hs = A @ A.T / d + 0.1 * np.eye(d)
        Hup = np.array([Hs[i, j] for j in range(d) for i in range(j + 1)])
        b = rng.standard_normal(d)
        print(f"newton d={d}", flush=True)
        ok, dx = E.newton_solve(Hup, d, 1e-3, b)
        ok_o, dx_o, _ = oracle.cholesky_solve(Hup, d, 1e-3, b)
        assert ok and ok_o
        np.testing.assert_allclose(dx, dx_o, rtol=1e-10, atol=1e-12 * np.abs(dx_o).max())


@pytest.mark.parametrize("d", [700, 1500, 2048])
def test_newton_solve_large_d_matches_oracle(oracle, d):
    """Block columns split into several 512-row segments (d > 512)."""
    from ufednl_amd import engine as E
    rng = np.random.default_rng(d)
    A = rng.standard_normal((d, d))
    Hs = A @ A.T / d + 0.1 * np.eye(d)
    Hup = np.concatenate([Hs[: j + 1, j] for j in range(d)])
    b = rng.standard_normal(d)
    print(f"newton d={d}", flush=True)
    ok, dx = E.newton_solve(Hup, d, 1e-3, b)
    ok_o, dx_o, _ = oracle.cholesky_solve(Hup, d, 1e-3, b)
    assert ok and ok_o
    np.testing.assert_allclose(dx, dx_o, rtol=1e-9, atol=1e-11 * np.abs(dx_o).max())
    # and the solve itself: (H + shift I) dx = b
    r = (Hs + 1e-3 * np.eye(d)) @ dx - b
    assert np.abs(r).max() <= 1e-10 * np.abs(b).max()


@pytest.mark.parametrize("d", [6, 300, 700])
def test_newton_solve_reports_failure_any_size(d):
    """A negative pivot in a late block column fails the factorization."""
    from ufednl_amd import engine as E
    Hs = np.eye(d)
    Hs[d - 1, d - 1] = -1.0
    Hup = np.concatenate([Hs[: j + 1, j] for j in range(d)])
    ok, dx = E.newton_solve(Hup, d, 0.0, np.ones(d))
    assert not ok
    assert np.all(dx == 0.0)


def test_newton_solve_reports_failure():
    from ufednl_amd import engine as E
    d = 6
    Hs = -np.eye(d)
    Hup = np.array([Hs[i, j] for j in range(d) for i in range(j + 1)])
    ok, _ = E.newton_solve(Hup, d, 0.0, np.ones(d))
    assert not ok
